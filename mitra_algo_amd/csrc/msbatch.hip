// Batched-seed mean-shift over compact votes (no reference counterpart: the
// reference only has sequential seeds, src/cpu/clustering.cpp:44-60).
//
// Points: rigid votes become (t_c, r) in R^6 -- translation about the model
// centre (length units) and the rotation's axis-angle vector (radians);
// reflection votes become (d_c, n) in R^4 -- plane offset about the centre
// and unit normal.  One bandwidth h applies to every coordinate.
// Kernel: Gaussian truncated at cutoff*h and shifted to reach zero there,
//   w = exp(-d^2 / 2h^2) - exp(-cutoff^2 / 2)   for d < cutoff*h,
// so the support boundary carries no weight jump.
//
// Grid: over the three coordinates of largest variance (for height fields:
// the two horizontal translation components and the rotation about z for
// rigid votes, plane offset and the horizontal normal components for mirror
// votes), cells of side >= cutoff*h / reach across the rows and
// cutoff*h / xdiv along them (first axis), so a cell holds few points
// outside the support: reach 3 and xdiv 3 for queries with a block per
// seed, reach 1 and xdiv 4 for the narrow ones, smaller when the grid would
// exceed its cell bound.  Points are radix-sorted by cell (stable: sums are
// deterministic) and indexed by per-cell offsets, each sorted point writing
// the run of cells that starts at it (no atomics).  With the first grid axis
// fastest, the cells of one row are contiguous, so a query visits (2 reach +
// 1)^2 contiguous point ranges (49 or 9), each cut to the x-cells its slice
// of the support ball reaches.  Queries run all iterations in registers and
// reduce with shuffles.  Seeds run in cell order and blocks are mapped so
// that each XCD takes a contiguous run of seeds: neighbouring queries share
// that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <limits>

#include "pairing.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, blocks dealt round-robin
constexpr int kMaxD = 6;

// Fused multiply-add, spelled out: the library builds without contraction
// (the pair and vote arithmetic must round as written), the mean-shift sums
// have no such constraint.
__device__ inline float ms_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double ms_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// exp(-d^2 / 2h^2) from d^2 and a folded scale c: FP32 as 2^(d^2 c) with
// c = -log2(e) / 2h^2 (one multiply and v_exp_f32), FP64 as exp(d^2 c) with
// c = -1 / 2h^2.  An infinite d^2 (the sentinel record) gives 0.
template <class R> __device__ inline R ms_gauss(R dd, R c);
template <> __device__ inline float ms_gauss<float>(float dd, float c) {
  return __builtin_amdgcn_exp2f(dd * c);
}
template <> __device__ inline double ms_gauss<double>(double dd, double c) { return exp(dd * c); }

template <class R>
__global__ void k_points_rigid(const R *q4, const R *t3, uint64_t n, R *pts) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  R w = q4[i], x = q4[n + i], y = q4[2 * n + i], z = q4[3 * n + i];
  R s = sqrt(x * x + y * y + z * z);
  R f = s > R(1e-12) ? R(2) * atan2(s, w) / s : R(2) / (w > R(0.5) ? w : R(1));
  pts[i] = t3[i];
  pts[n + i] = t3[n + i];
  pts[2 * n + i] = t3[2 * n + i];
  pts[3 * n + i] = f * x;
  pts[4 * n + i] = f * y;
  pts[5 * n + i] = f * z;
}

// Mirror normals come as 3 rows of stride `stride` (the emission capacity).
template <class R>
__global__ void k_points_refl(const R *n3, uint64_t stride, const R *d1, uint64_t n, R *pts) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  pts[i] = d1[i];
  pts[n + i] = n3[i];
  pts[2 * n + i] = n3[stride + i];
  pts[3 * n + i] = n3[2 * stride + i];
}

// Per-block box and moments of every coordinate: partial[b * kPart + k]
// holds min (k < kMaxD), max, sum and sum of squares (FP64) in blocks of
// kMaxD.
constexpr int kPart = 4 * kMaxD;
template <class R, int D>
__global__ void __launch_bounds__(kThreads) k_bbox(const R *pts, uint64_t n, double *partial) {
  __shared__ double sh[kWavesPerBlock][kPart];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  double v[kPart];
  for (int k = 0; k < kMaxD; ++k)
    v[k] = INFINITY, v[kMaxD + k] = -INFINITY, v[2 * kMaxD + k] = 0, v[3 * kMaxD + k] = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads)
    for (int k = 0; k < D; ++k) {
      const double x = static_cast<double>(pts[k * n + i]);
      v[k] = fmin(v[k], x);
      v[kMaxD + k] = fmax(v[kMaxD + k], x);
      v[2 * kMaxD + k] += x;
      v[3 * kMaxD + k] += x * x;
    }
  for (int o = kWave / 2; o > 0; o >>= 1)
    for (int k = 0; k < kMaxD; ++k) {
      v[k] = fmin(v[k], __shfl_xor(v[k], o, kWave));
      v[kMaxD + k] = fmax(v[kMaxD + k], __shfl_xor(v[kMaxD + k], o, kWave));
      v[2 * kMaxD + k] += __shfl_xor(v[2 * kMaxD + k], o, kWave);
      v[3 * kMaxD + k] += __shfl_xor(v[3 * kMaxD + k], o, kWave);
    }
  if (lane == 0)
    for (int k = 0; k < kPart; ++k)
      sh[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kPart) {
    const int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kWavesPerBlock; ++w) {
      const double x = sh[w][k];
      r = k < kMaxD ? fmin(r, x) : (k < 2 * kMaxD ? fmax(r, x) : r + x);
    }
    partial[blockIdx.x * kPart + k] = r;
  }
}

// Segments (batches of meshes): item i of a point or seed list belongs to
// the last m with off[m] <= i (off: meshes + 1 starts); off == nullptr means
// one segment.
__device__ inline uint32_t seg_of(const uint64_t *off, uint32_t meshes, uint64_t i) {
  if (!off)
    return 0;
  uint32_t f = 0, c = meshes + 1;
  while (c) {
    const uint32_t h = c / 2, m = f + h;
    if (off[m] <= i) {
      f = m + 1;
      c -= h + 1;
    } else
      c = h;
  }
  return f - 1;
}

// Box and moments of the segments (layout of k_bbox): kSegBlocks blocks per
// segment, block p of segment m taking every kSegBlocks-th run of kThreads
// points, so one large mesh among small ones does not hold the launch.
constexpr unsigned kSegBlocks = 8;
template <class R, int D>
__global__ void __launch_bounds__(kThreads)
    k_seg_moments(const R *pts, uint64_t n, const uint64_t *seg_off, double *partial) {
  __shared__ double sh[kWavesPerBlock][kPart];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint32_t mesh = blockIdx.x / kSegBlocks, part = blockIdx.x % kSegBlocks;
  const uint64_t b = seg_off[mesh], e = seg_off[mesh + 1];
  double v[kPart];
  for (int k = 0; k < kMaxD; ++k)
    v[k] = INFINITY, v[kMaxD + k] = -INFINITY, v[2 * kMaxD + k] = 0, v[3 * kMaxD + k] = 0;
  for (uint64_t i = b + static_cast<uint64_t>(part) * kThreads + threadIdx.x; i < e;
       i += static_cast<uint64_t>(kSegBlocks) * kThreads)
    for (int k = 0; k < D; ++k) {
      const double x = static_cast<double>(pts[k * n + i]);
      v[k] = fmin(v[k], x);
      v[kMaxD + k] = fmax(v[kMaxD + k], x);
      v[2 * kMaxD + k] += x;
      v[3 * kMaxD + k] += x * x;
    }
  for (int o = kWave / 2; o > 0; o >>= 1)
    for (int k = 0; k < kMaxD; ++k) {
      v[k] = fmin(v[k], __shfl_xor(v[k], o, kWave));
      v[kMaxD + k] = fmax(v[kMaxD + k], __shfl_xor(v[kMaxD + k], o, kWave));
      v[2 * kMaxD + k] += __shfl_xor(v[2 * kMaxD + k], o, kWave);
      v[3 * kMaxD + k] += __shfl_xor(v[3 * kMaxD + k], o, kWave);
    }
  if (lane == 0)
    for (int k = 0; k < kPart; ++k)
      sh[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kPart) {
    const int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kWavesPerBlock; ++w) {
      const double x = sh[w][k];
      r = k < kMaxD ? fmin(r, x) : (k < 2 * kMaxD ? fmax(r, x) : r + x);
    }
    partial[static_cast<uint64_t>(blockIdx.x) * kPart + k] = r;
  }
}

// perm lists the point coordinates in storage (slot) order: the three grid
// axes ascending, then the remaining coordinates ascending.  Rows run along
// grid axis 0 (x): cells have side cellx along it and side cell along axes 1
// and 2, with cell >= cutoff*h / reach, so a query visits the (2 reach + 1)^2
// rows around its (y, z) cell, each cut to the x-cells its slice of the
// support ball reaches.
struct Grid {
  int axis[3];
  int perm[kMaxD];
  double lo[3];
  double cell;   // side along axes 1 and 2
  double inv;    // 1 / cell: cell_of multiplies, in the keys and the queries alike
  double cellx;  // side along axis 0 (the rows)
  double invx;   // 1 / cellx
  int dims[3];
  int reach;
  uint64_t ncells;
};
constexpr int kMaxReach = 3;  // 7 x 7 = 49 rows: one row range per lane
// Grids of the narrow queries (many seeds, few candidates each): rows one
// cell of side cutoff*h to either side (3 x 3 rows, so few short batches),
// cells a quarter of that along the rows (a tight cut to the ball).
constexpr int kNarrowReach = 1;
constexpr int kNarrowXdiv = 4;
constexpr uint32_t kMapPerWave = 1024;  // LDS candidate-map entries per wave
constexpr int kMaxRows = (2 * kMaxReach + 1) * (2 * kMaxReach + 1);
constexpr int kCellBits = 22;  // at most 2^22 cells: three 8-bit radix passes
constexpr uint64_t kMaxCells = 1ULL << kCellBits;
constexpr unsigned kBboxBlocks = 1024;
// points per thread and round of k_points_moments (FP64: 2, for registers)
template <class R> constexpr int kPointsPerRound = sizeof(R) == 4 ? 4 : 2;

// Device-side control block of one mean-shift call: the grid, the sums of
// up to two runs (run k: candidates, in-support visits, seed-iterations at
// sums[3k..3k+2]) and the non-finite-input flag, read back once at the end.
struct MsCtl {
  Grid grid;
  unsigned long long sums[kMsSums];
  unsigned bad;
  Grid grid2;  // the second slot set's grid
};

// One block per segment: reduces the segment's nb partial boxes and moments
// (lanes over partials, then the four waves through LDS) and lays its grid
// on the three coordinates of largest spread (variance; ties: lower index
// first).  Spread, not extent: a few outlying votes can stretch a coordinate
// whose votes are otherwise concentrated.
// The first (reach, x subdivision) pair whose grid has at most
// clamp(2 n, 1024, kMaxCells) cells wins: reach from max_reach down to 1,
// and for each reach the x subdivision from max(xdiv, reach) down to reach
// (cells of side cell0 / reach across the rows, cell0 / xdiv along them);
// if none fits, cells of side cell0, doubled until the grid fits.
// No points: a single cell.  ncells[m] (optional) receives the cell count.
// (One block: the segment's nb partials from `partial`, its n points; the
// grid goes to *gout, the cell count to *ncells when given.)
template <int D>
__device__ void grid_block(const double *partial, unsigned nb, uint64_t n, double cell0,
                           int max_reach, int xdiv, Grid *gout, uint64_t *ncells,
                           unsigned *bad) {
  __shared__ double red[kWavesPerBlock][4 * kMaxD];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t cap = n < kMaxCells / 2 ? (2 * n > 1024 ? 2 * n : 1024) : kMaxCells;
  double lo[kMaxD], hi[kMaxD], sum[kMaxD], sq[kMaxD];
  for (int k = 0; k < kMaxD; ++k)
    lo[k] = INFINITY, hi[k] = -INFINITY, sum[k] = 0, sq[k] = 0;
  for (unsigned b = threadIdx.x; b < nb; b += kThreads)
    for (int k = 0; k < D; ++k) {
      lo[k] = fmin(lo[k], partial[b * kPart + k]);
      hi[k] = fmax(hi[k], partial[b * kPart + kMaxD + k]);
      sum[k] += partial[b * kPart + 2 * kMaxD + k];
      sq[k] += partial[b * kPart + 3 * kMaxD + k];
    }
  for (int k = 0; k < D; ++k)
    for (int o = kWave / 2; o > 0; o >>= 1) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, kWave));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, kWave));
      sum[k] += __shfl_xor(sum[k], o, kWave);
      sq[k] += __shfl_xor(sq[k], o, kWave);
    }
  if (lane == 0)
    for (int k = 0; k < D; ++k) {
      red[wave][k] = lo[k];
      red[wave][kMaxD + k] = hi[k];
      red[wave][2 * kMaxD + k] = sum[k];
      red[wave][3 * kMaxD + k] = sq[k];
    }
  __syncthreads();
  if (threadIdx.x != 0)
    return;
  for (int w = 1; w < kWavesPerBlock; ++w)
    for (int k = 0; k < D; ++k) {
      lo[k] = fmin(lo[k], red[w][k]);
      hi[k] = fmax(hi[k], red[w][kMaxD + k]);
      sum[k] += red[w][2 * kMaxD + k];
      sq[k] += red[w][3 * kMaxD + k];
    }
  double var[kMaxD];
  for (int k = 0; k < D; ++k) {
    const double mean = n ? sum[k] / static_cast<double>(n) : 0.0;
    var[k] = n ? sq[k] / static_cast<double>(n) - mean * mean : 0.0;
  }
  Grid g{};
  g.axis[0] = 0, g.axis[1] = 1, g.axis[2] = 2;
  for (int k = 0; k < kMaxD; ++k)
    g.perm[k] = k;
  g.cell = cell0;
  g.inv = 1.0 / cell0;
  g.cellx = cell0;
  g.invx = 1.0 / cell0;
  g.dims[0] = g.dims[1] = g.dims[2] = 1;
  g.reach = 1;
  g.ncells = 1;
  bool finite = true;
  for (int k = 0; k < D; ++k)
    finite = finite && isfinite(lo[k]) && isfinite(hi[k]);
  if (nb && n && !finite)
    atomicOr(bad, 1u);
  if (nb && n && finite) {
    // the three largest variances (ties: lower index first), as a bit mask;
    // fixed-trip loops only, so everything stays in registers
    unsigned sel = 0;
    for (int k = 0; k < D; ++k) {
      int rank = 0;
      for (int j = 0; j < D; ++j)
        rank += var[j] > var[k] || (var[j] == var[k] && j < k);
      sel |= rank < 3 ? 1u << k : 0u;
    }
    unsigned rest = ((1u << D) - 1) & ~sel;
    int perm[kMaxD] = {0, 1, 2, 3, 4, 5};
    for (int k = 0; k < 3; ++k) {
      perm[k] = __builtin_ctz(sel);
      sel &= sel - 1;
    }
    for (int k = 3; k < D; ++k) {
      perm[k] = __builtin_ctz(rest);
      rest &= rest - 1;
    }
    double alo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
      alo[k] = 0, ext[k] = 0;
      for (int j = 0; j < D; ++j)
        if (perm[k] == j)
          alo[k] = lo[j], ext[k] = hi[j] - lo[j];
    }
    // cells of side cx along the rows and cyz across them (see above)
    auto fits = [&](double cx, double cyz) {
      uint64_t cells = 1;
      for (int k = 0; k < 3; ++k) {
        const double d = floor(ext[k] * (1.0 / (k == 0 ? cx : cyz))) + 1;
        g.dims[k] = d < 2147483647.0 ? static_cast<int>(d) : 2147483647;
        cells = cells > cap ? cells : cells * static_cast<uint64_t>(g.dims[k]);
      }
      g.ncells = cells;
      return cells <= cap;
    };
    int reach = 1;
    double cell = cell0, cellx = cell0;
    bool found = false;
    for (int r = max_reach; r >= 1 && !found; --r)
      for (int xd = xdiv > r ? xdiv : r; xd >= r && !found; --xd)
        if (fits(cell0 / xd, cell0 / r)) {
          found = true;
          reach = r;
          cell = cell0 / r;
          cellx = cell0 / xd;
        }
    if (!found) {
      while (!fits(cell, cell))
        cell *= 2;
      cellx = cell;
    }
    g.reach = reach;
    for (int k = 0; k < 3; ++k) {
      g.axis[k] = perm[k];
      g.lo[k] = alo[k];
    }
    for (int k = 0; k < kMaxD; ++k)
      g.perm[k] = perm[k];
    g.cell = cell;
    g.inv = 1.0 / cell;
    g.cellx = cellx;
    g.invx = 1.0 / cellx;
  }
  *gout = g;
  if (ncells)
    *ncells = g.ncells;
}
template <int D>
__global__ void __launch_bounds__(kThreads)
    k_grid_setup(const double *partial, unsigned nb, const uint64_t *seg_off, uint64_t n,
                 double cell0, int max_reach, int xdiv, Grid *grids, uint64_t *ncells,
                 unsigned *bad) {
  const uint32_t mesh = blockIdx.x;
  if (seg_off)
    n = seg_off[mesh + 1] - seg_off[mesh];
  grid_block<D>(partial + static_cast<uint64_t>(mesh) * nb * kPart, nb, n, cell0, max_reach,
                xdiv, grids + mesh, ncells ? ncells + mesh : nullptr, bad);
}

// The points of one kind and, in the same launch, their box and moments.
// Rigid (D = 6): a = q4, b = t3 (rows of n), as k_points_rigid; mirror (D =
// 4): a = n3 (rows of `stride`), b = d1, as k_points_refl.  Each block
// reduces its points' box and moments as k_bbox does into partial[block]
// (saving the box launch and its re-read of the points).  A last-block grid
// in the same launch was slower: the device-scope fence each block needs
// before its ticket writes the die's L2 back, full of the points just
// written.
template <class R, int D>
__global__ void __launch_bounds__(kThreads)
    k_points_moments(const R *a, const R *b, uint64_t stride, uint64_t n, R *pts,
                     double *partial) {
  __shared__ double sh[kWavesPerBlock][kPart];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  // box in the points' own type (exact), sums and sums of squares in FP64
  R lo[D], hi[D];
  double sm[D], sq[D];
  for (int k = 0; k < D; ++k)
    lo[k] = R(INFINITY), hi[k] = R(-INFINITY), sm[k] = 0, sq[k] = 0;
  // kPointsPerRound<R> points per thread and round, all their loads in flight
  // before the arithmetic (rigid: q4 then t3; mirror: d then n3)
  constexpr int kIn = D == 6 ? 7 : 4;
  const uint64_t step = static_cast<uint64_t>(gridDim.x) * kThreads;
  for (uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i0 < n;
       i0 += kPointsPerRound<R> * step) {
    R in[kPointsPerRound<R>][kIn];
    for (int u = 0; u < kPointsPerRound<R>; ++u) {
      const uint64_t i = i0 + u * step;
      if (i >= n)
        continue;
      // 32-bit lane offsets from per-row scalar bases (n < 2^32)
      const uint32_t iu = static_cast<uint32_t>(i);
      if constexpr (D == 6) {
        for (int k = 0; k < 4; ++k)
          in[u][k] = (a + k * n)[iu];
        for (int k = 0; k < 3; ++k)
          in[u][4 + k] = (b + k * n)[iu];
      } else {
        in[u][0] = b[iu];
        for (int k = 0; k < 3; ++k)
          in[u][1 + k] = (a + k * stride)[iu];
      }
    }
    for (int u = 0; u < kPointsPerRound<R>; ++u) {
      const uint64_t i = i0 + u * step;
      if (i >= n)
        continue;
      R c[D];
      if constexpr (D == 6) {
        const R w = in[u][0], x = in[u][1], y = in[u][2], z = in[u][3];
        const R sn = sqrt(x * x + y * y + z * z);
        const R f = sn > R(1e-12) ? R(2) * atan2(sn, w) / sn : R(2) / (w > R(0.5) ? w : R(1));
        c[0] = in[u][4], c[1] = in[u][5], c[2] = in[u][6];
        c[3] = f * x, c[4 % D] = f * y, c[5 % D] = f * z;
      } else {
        for (int k = 0; k < D; ++k)
          c[k] = in[u][k];
      }
      for (int k = 0; k < D; ++k) {
        (pts + k * n)[static_cast<uint32_t>(i)] = c[k];
        const double x = static_cast<double>(c[k]);
        lo[k] = fmin(lo[k], c[k]);
        hi[k] = fmax(hi[k], c[k]);
        sm[k] += x;
        sq[k] += x * x;
      }
    }
  }
  for (int o = kWave / 2; o > 0; o >>= 1)
    for (int k = 0; k < D; ++k) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, kWave));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, kWave));
      sm[k] += __shfl_xor(sm[k], o, kWave);
      sq[k] += __shfl_xor(sq[k], o, kWave);
    }
  if (lane == 0)
    for (int k = 0; k < kMaxD; ++k) {
      const bool in = k < D;
      sh[wave][k] = in ? static_cast<double>(lo[k % D]) : INFINITY;
      sh[wave][kMaxD + k] = in ? static_cast<double>(hi[k % D]) : -INFINITY;
      sh[wave][2 * kMaxD + k] = in ? sm[k % D] : 0.0;
      sh[wave][3 * kMaxD + k] = in ? sq[k % D] : 0.0;
    }
  __syncthreads();
  if (threadIdx.x < kPart) {
    const int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kWavesPerBlock; ++w) {
      const double x = sh[w][k];
      r = k < kMaxD ? fmin(r, x) : (k < 2 * kMaxD ? fmax(r, x) : r + x);
    }
    partial[static_cast<uint64_t>(blockIdx.x) * kPart + k] = r;
  }
}

__device__ inline uint64_t lane_u64(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), l);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), l);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Cell index along one axis (inv = 1 / cell side); NaN and underflow go to
// cell 0.  Keys and queries use this one formula, so they agree on every
// point's cell; cells are monotone in v.
__device__ inline int cell_of(double v, double lo, double inv, int dim) {
  double c = floor((v - lo) * inv);
  if (!(c > 0))
    return 0;
  return c >= dim ? dim - 1 : static_cast<int>(c);
}
template <class R> __device__ inline uint64_t cell_key(const Grid &g, R a, R b, R c) {
  int x = cell_of(static_cast<double>(a), g.lo[0], g.invx, g.dims[0]);
  int y = cell_of(static_cast<double>(b), g.lo[1], g.inv, g.dims[1]);
  int z = cell_of(static_cast<double>(c), g.lo[2], g.inv, g.dims[2]);
  return (static_cast<uint64_t>(z) * g.dims[1] + y) * g.dims[0] + x;
}

// Neighbourhood rows.  Row k (0 <= k < (2 reach + 1)^2: dz = k / span -
// reach, dy = k % span - reach around the seed's (y, z) cell (cy, cz)) is
// cut to the x-cells its slice of the support ball can reach: a point of row
// (y, z) is at least ey, ez from the seed in those two coordinates, so within
// the support only if |x - m0| <= sqrt(rlim^2 - ey^2 - ez^2).  Cells are
// monotone in each coordinate (also where clamped at the grid edges); rlim
// and the slab bounds (slop) carry margins far above the rounding of the
// FP32 / FP64 distance test, the half-width has a relative margin of 1e-6
// over its FP32 root and the x-cells one of 1e-6 cells over the product by
// 1 / cellx: no point inside the support is cut.  Returns the row's first
// sorted record and its length (0: the row misses the ball or the grid).
struct RowCut {
  double slop, rlim2;
};
__device__ inline RowCut row_cut(const Grid &g, double r2) {
  RowCut c;
  c.slop = 1e-9 * (1.0 + fabs(g.lo[1]) + fabs(g.lo[2]) + g.cell * (g.dims[1] + g.dims[2]));
  const double rlim = sqrt(r2) * (1.0 + 1e-5) + c.slop;
  c.rlim2 = rlim * rlim;
  return c;
}
__device__ inline void cut_row(const Grid &g, const RowCut &c, uint64_t cb,
                               const uint64_t *offsets, double mx, double my, double mz, int cy,
                               int cz, int k, uint32_t &rb, uint32_t &len) {
  rb = 0, len = 0;
  const int span = 2 * g.reach + 1;
  const int z = cz + k / span - g.reach, y = cy + k % span - g.reach;
  if (k >= span * span || z < 0 || z >= g.dims[2] || y < 0 || y >= g.dims[1])
    return;
  double ey = y > cy ? g.lo[1] + y * g.cell - my : (y < cy ? my - (g.lo[1] + (y + 1) * g.cell) : 0.0);
  double ez = z > cz ? g.lo[2] + z * g.cell - mz : (z < cz ? mz - (g.lo[2] + (z + 1) * g.cell) : 0.0);
  ey = ey > c.slop ? ey - c.slop : 0.0;
  ez = ez > c.slop ? ez - c.slop : 0.0;
  const double rem = c.rlim2 - ey * ey - ez * ez;
  if (!(rem > 0))
    return;
  const double hx = static_cast<double>(sqrtf(static_cast<float>(rem))) * (1.0 + 1e-6);
  const double ua = floor((mx - hx - g.lo[0]) * g.invx - 1e-6);
  const double ub = floor((mx + hx - g.lo[0]) * g.invx + 1e-6);
  const int last = g.dims[0] - 1;
  const int xa = ua > 0 ? (ua < last + 1 ? static_cast<int>(ua) : last + 1) : 0;
  const int xb = ub < last ? (ub > -1 ? static_cast<int>(ub) : -1) : last;
  if (xa > xb)
    return;
  const uint64_t row = cb + (static_cast<uint64_t>(z) * g.dims[1] + y) * g.dims[0];
  const uint64_t b = offsets[row + xa];
  rb = static_cast<uint32_t>(b);
  len = static_cast<uint32_t>(offsets[row + xb + 1] - b);
}

// Storage of the cell-sorted points: slot coordinates 0-3 (three grid axes
// and the first remaining one) as one 4-vector, slots 4-5 (6-D only) as a
// 2-vector, so the three-axis pretest is one coalesced load per candidate.
template <class R> struct Vec;
template <> struct Vec<float> {
  using V4 = float4;
  using V2 = float2;
};
template <> struct Vec<double> {
  using V4 = double4;
  using V2 = double2;
};

// Global cell id = the segment's cell base + its local cell id, so the
// sorted order is segment-major.  Points past the last segment (a graph
// replay whose inputs gave fewer votes than the captured step) take the key
// `beyond`, larger than every cell id: they sort last and no cell holds them.
// The pass also copies each point into `aos` (kAos scalars per point,
// coordinates in input order), so the gather after the sort reads one
// 32- or 64-byte record per point instead of D scattered scalars.
constexpr int kAos = 8;
template <class R, int D>
__global__ void k_cell_keys(const R *pts, uint64_t n, const Grid *grids, const uint64_t *seg_off,
                            uint32_t meshes, const uint64_t *cellbase, uint64_t beyond,
                            uint64_t *keys, uint32_t *vals, R *aos) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  R v[kAos];
  for (int k = 0; k < kAos; ++k)
    v[k] = k < D ? pts[k * n + i] : R(0);
  for (int k = 0; k < kAos; k += 2) {  // 8- or 16-byte stores, record-contiguous
    typename Vec<R>::V2 u;
    u.x = v[k], u.y = v[k + 1];
    *reinterpret_cast<typename Vec<R>::V2 *>(aos + i * kAos + k) = u;
  }
  const uint32_t m = seg_of(seg_off, meshes, i);
  vals[i] = static_cast<uint32_t>(i);
  if (m >= meshes) {
    keys[i] = beyond;
    return;
  }
  const Grid g = grids[m];
  R a = v[0], b = v[0], c = v[0];
  for (int k = 1; k < D; ++k) {  // the grid axes' coordinates, without dynamic indexing
    a = g.axis[0] == k ? v[k] : a;
    b = g.axis[1] == k ? v[k] : b;
    c = g.axis[2] == k ? v[k] : c;
  }
  keys[i] = (cellbase ? cellbase[m] : 0) + cell_key(g, a, b, c);
}

// offsets[c] = the first sorted record whose key is >= c, c in [0, ncells]:
// record i (0 <= i <= n) starts the cells (key[i - 1], key[i]] (key[-1] = -1,
// key[n] = ncells; keys past ncells, the segments' sentinels, clamp), so each
// record writes its run of cells.  Runs of up to kShortRun cells are written
// by their lane, longer runs (empty stretches of the grid) by the whole wave.
constexpr uint64_t kShortRun = 8;
// (block blk of nblk)
__device__ inline void cell_offsets_block(unsigned blk, unsigned nblk, const uint64_t *skeys,
                                          uint64_t n, const uint64_t *total, uint64_t *offsets) {
  const uint64_t ncells = *total;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t stride = static_cast<uint64_t>(nblk) * kThreads;
  for (uint64_t base = static_cast<uint64_t>(blk) * kThreads + (threadIdx.x - lane);
       base <= n; base += stride) {
    const uint64_t i = base + static_cast<uint64_t>(lane);
    uint64_t lo = 0, hi = 0;  // the run: cells [lo, hi)
    if (i <= n) {
      lo = i == 0 ? 0 : min(skeys[i - 1] + 1, ncells + 1);
      hi = i == n ? ncells + 1 : min(skeys[i] + 1, ncells + 1);
      hi = hi > lo ? hi : lo;
    }
    const bool shortrun = hi - lo <= kShortRun;
    if (shortrun)
      for (uint64_t c = lo; c < hi; ++c)
        offsets[c] = i;
    uint64_t longer = ballot(!shortrun);
    while (longer) {
      const int l = __builtin_ctzll(longer);
      longer &= longer - 1;
      const uint64_t rlo = lane_u64(lo, l), rhi = lane_u64(hi, l);
      for (uint64_t c = rlo + static_cast<uint64_t>(lane); c < rhi; c += kWave)
        offsets[c] = base + static_cast<uint64_t>(l);
    }
  }
}

__global__ void __launch_bounds__(kThreads)
    k_cell_offsets(const uint64_t *skeys, uint64_t n, const uint64_t *total,
                   uint64_t *offsets) {
  cell_offsets_block(blockIdx.x, gridDim.x, skeys, n, total, offsets);
}

template <class R, int D>
__device__ inline void gather_sorted_block(unsigned blk, const R *aos, uint64_t n,
                                           const Grid *grids, const uint64_t *seg_off,
                                           uint32_t meshes, const uint32_t *order,
                                           typename Vec<R>::V4 *p4, typename Vec<R>::V2 *q2) {
  uint64_t i = static_cast<uint64_t>(blk) * kThreads + threadIdx.x;
  if (i >= n)
    return;
  if (i == 0) {  // the sentinel record at [n]: outside every support
    typename Vec<R>::V4 far;
    far.x = far.y = far.z = far.w = R(1e30);
    p4[n] = far;
    if (D == 6) {
      typename Vec<R>::V2 farq;
      farq.x = farq.y = R(1e30);
      q2[n] = farq;
    }
  }
  const uint64_t o = order[i];
  const uint32_t m = seg_of(seg_off, meshes, o);
  if (m >= meshes)
    return;  // past the last segment: in no cell, never read
  const int *perm = grids[m].perm;
  R v[kAos];
  for (int k = 0; k < kAos; k += 2) {
    const typename Vec<R>::V2 u = *reinterpret_cast<const typename Vec<R>::V2 *>(aos + o * kAos + k);
    v[k] = u.x, v[k + 1] = u.y;
  }
  // slot j holds coordinate perm[j] (selects, no dynamic register indexing)
  R w[D];
  for (int j = 0; j < D; ++j) {
    const int pj = perm[j];
    R x = v[0];
    for (int k = 1; k < D; ++k)
      x = pj == k ? v[k] : x;
    w[j] = x;
  }
  typename Vec<R>::V4 a;
  a.x = w[0], a.y = w[1], a.z = w[2], a.w = w[3];
  p4[i] = a;
  if (D == 6) {
    typename Vec<R>::V2 u;
    u.x = w[4 % D], u.y = w[5 % D];
    q2[i] = u;
  }
}

// Seeds taken straight from the cell-sorted points: seed j of the k seeds of
// a segment of n points is drawn from the j-th of k equal strata of the
// segment's sorted range (splitmix64 of (salt, j)), so seeds come out
// distinct and in cell order.  ids receive the seeds' point indices in the
// caller's (unsorted) numbering, seed_mesh their segment.  Segments: seed_off
// (meshes + 1) splits the seeds, seg_off the points; one segment otherwise.
__device__ inline void pick_sorted_block(unsigned blk, const uint64_t *seg_off,
                                         const uint64_t *seed_off, uint32_t meshes, uint64_t n,
                                         uint64_t k, uint64_t salt, const uint32_t *sorted_ids,
                                         uint32_t *pos, int32_t *ids, uint32_t *seed_mesh) {
  const uint64_t j = static_cast<uint64_t>(blk) * kThreads + threadIdx.x;
  if (j >= (seed_off ? seed_off[meshes] : k))
    return;
  const uint32_t m = seg_of(seed_off, meshes, j);
  const uint64_t jl = seed_off ? j - seed_off[m] : j;
  const uint64_t km = seed_off ? seed_off[m + 1] - seed_off[m] : k;
  const uint64_t base = seg_off ? seg_off[m] : 0;
  const uint64_t nm = seg_off ? seg_off[m + 1] - base : n;
  const uint64_t lo = jl * nm / km, hi = (jl + 1) * nm / km;
  const uint64_t p = base + lo + splitmix64(salt ^ splitmix64(jl)) % (hi - lo);
  pos[j] = static_cast<uint32_t>(p);
  ids[j] = static_cast<int32_t>(sorted_ids[p]);
  if (seed_mesh)
    seed_mesh[j] = m;
}

__global__ void __launch_bounds__(kThreads)
    k_pick_sorted(const uint64_t *seg_off, const uint64_t *seed_off, uint32_t meshes, uint64_t n,
                  uint64_t k, uint64_t salt, const uint32_t *sorted_ids, uint32_t *pos,
                  int32_t *ids, uint32_t *seed_mesh) {
  pick_sorted_block(blockIdx.x, seg_off, seed_off, meshes, n, k, salt, sorted_ids, pos, ids,
                    seed_mesh);
}

// What follows the cell sort, in one launch of three block ranges that read
// the sorted keys and order only: the cell offsets (nb_off blocks), the
// gather of the sorted records (nb_gather), and the drawn seeds (the rest;
// none without seed_ids).
struct SortedArgs {
  const uint64_t *skeys;
  const uint64_t *total;
  uint64_t *offsets;
  const uint64_t *seg_off, *seed_off;
  uint32_t meshes;
  uint64_t n, ns, salt;
  const uint32_t *sorted_ids;
  uint32_t *pos;
  int32_t *ids;
  uint32_t *seed_mesh;
  const Grid *grids;
};
template <class R, int D>
__global__ void __launch_bounds__(kThreads)
    k_after_sort(unsigned nb_off, unsigned nb_gather, SortedArgs A, const R *aos,
                 typename Vec<R>::V4 *p4, typename Vec<R>::V2 *q2) {
  const unsigned b = blockIdx.x;
  if (b < nb_off)
    cell_offsets_block(b, nb_off, A.skeys, A.n, A.total, A.offsets);
  else if (b < nb_off + nb_gather)
    gather_sorted_block<R, D>(b - nb_off, aos, A.n, A.grids, A.seg_off, A.meshes, A.sorted_ids,
                              p4, q2);
  else
    pick_sorted_block(b - nb_off - nb_gather, A.seg_off, A.seed_off, A.meshes, A.n, A.ns, A.salt,
                      A.sorted_ids, A.pos, A.ids, A.seed_mesh);
}

// Explicit seed positions -> cell keys (to run those seeds in cell order).
template <class R>
__global__ void k_seed_keys(const R *seeds, uint64_t ns, const Grid *grids, uint64_t *keys,
                            uint32_t *vals) {
  uint64_t s = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= ns)
    return;
  const Grid g = grids[0];
  keys[s] = cell_key(g, seeds[g.axis[0] * ns + s], seeds[g.axis[1] * ns + s],
                     seeds[g.axis[2] * ns + s]);
  vals[s] = static_cast<uint32_t>(s);
}


// Mean-shift queries.  A seed is served by W wavefronts (W = 1: four seeds
// per 256-lane block; W = 4: one seed per block, for few seeds with many
// candidates).  Seeds are sorted points (seed_pos, already in cell order) or
// explicit positions (seeds, D x ns SoA, run in `order`).  All iterations
// run in registers.  Per iteration the rows of the neighbourhood (each row
// 2 reach + 1 cells, contiguous in the sorted order) are flattened into one
// candidate range: a map in LDS sends each flattened candidate to its
// record, and the waves walk the range in 64-candidate batches with the next
// batches' records in flight (batch k goes to wave k mod W).  With W = 4 the
// first wave alone finds the rows (the mean's cell, the row ranges cut to the
// support ball, their prefix sum) and leaves them in LDS; each wave fills
// the map for its share of the rows; after the walk the waves' partial sums
// meet in LDS lane by lane in a fixed order, the first wave reduces them and
// updates the mean, and the others read the new mean back.  A seed whose
// update leaves every coordinate unchanged has reached a fixed point of the
// iteration and stops (further iterations would repeat it bit for bit).
template <class R, int D, int W>
__global__ void __launch_bounds__(kThreads)
    k_ms_query(const typename Vec<R>::V4 *p4, const typename Vec<R>::V2 *q2,
               const uint64_t *offsets, const Grid *grids, const uint64_t *cellbase,
               const uint32_t *seed_mesh, const uint64_t *ns_dev, const uint32_t *seed_pos,
               const R *seeds, const uint32_t *order, uint64_t ns, unsigned logical_blocks,
               uint32_t sentinel, int iterations, R gscale, R r2, R floor_w, R *out, R *weight,
               unsigned long long *stats) {
  static_assert(W == 1 || W == kWavesPerBlock, "one seed per wave or per block");
  // W = 1: candidate -> record map, one slice per wave.  W = 4: the
  // non-empty rows of this iteration (first record, length, first batch;
  // entry [nrc] of the batch starts holds the total), the waves' lane-wise
  // partial sums, the new mean with its weight, and the stop flag
  constexpr uint32_t kMapCap = kMapPerWave;
  constexpr int kRowSlots = W == 1 ? 1 : kMaxRows + 1;
  __shared__ uint32_t cmap[W == 1 ? kMapPerWave * kWavesPerBlock : 1];
  __shared__ uint32_t s_rb[kRowSlots], s_len[kRowSlots], s_bstart[kRowSlots];
  __shared__ uint32_t s_count[3];  // W = 4: non-empty rows, batches, candidates
  __shared__ R red[W == 1 ? 1 : (D + 1) * kThreads];
  __shared__ R s_mean[W == 1 ? 1 : D + 1];
  const int lane = threadIdx.x & (kWave - 1);
  // wave index, made visibly wave-uniform so the row cursor stays scalar
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int wsub = W == 1 ? 0 : wave;  // this wave's share of the seed's batches
  uint32_t *map = cmap + (W == 1 ? wave * kMapPerWave : 0);
  // XCD-aware block order: dies take contiguous runs of (cell-ordered) seeds
  const unsigned per = (logical_blocks + kXcds - 1) / kXcds;
  const unsigned logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (logical >= logical_blocks)
    return;
  const uint64_t slot = W == 1 ? static_cast<uint64_t>(logical) * kWavesPerBlock + wave
                               : static_cast<uint64_t>(logical);
  if (slot >= (ns_dev ? *ns_dev : ns))
    return;  // W = 4: the whole block leaves together
  const uint64_t s = order ? order[slot] : slot;
  // the seed's segment: its grid, and its cells' base in the offsets
  const uint32_t gm = seed_mesh ? seed_mesh[s] : 0;
  const Grid g = grids[gm];
  const uint64_t cb = cellbase ? cellbase[gm] : 0;
  R m[D];  // slot order
  if (seed_pos) {
    const uint64_t p = seed_pos[s];
    const typename Vec<R>::V4 v = p4[p];
    m[0] = v.x, m[1] = v.y, m[2] = v.z, m[3] = v.w;
    if (D == 6) {
      const typename Vec<R>::V2 u = q2[p];
      m[4 % D] = u.x, m[5 % D] = u.y;
    }
  } else {
    for (int k = 0; k < D; ++k)
      m[k] = seeds[g.perm[k] * ns + s];
  }
  R wsum = 0;
  unsigned long long cand = 0, inside = 0;
  int done = 0;  // updates computed
  const int span = 2 * g.reach + 1, nrows = span * span;
  const RowCut rc = row_cut(g, static_cast<double>(r2));
  for (int it = 0; it < iterations; ++it) {
    // lanes 0..nrows-1 fetch the ranges of the neighbourhood rows, each cut
    // to the support ball (cut_row)
    uint32_t rb = 0, startv = 0, rend = 0, total = 0;
    if (W == 1 || wave == 0) {
      const int cy = cell_of(static_cast<double>(m[1]), g.lo[1], g.inv, g.dims[1]);
      const int cz = cell_of(static_cast<double>(m[2]), g.lo[2], g.inv, g.dims[2]);
      uint32_t len = 0;
      if (lane < nrows)
        cut_row(g, rc, cb, offsets, static_cast<double>(m[0]), static_cast<double>(m[1]),
                static_cast<double>(m[2]), cy, cz, lane, rb, len);
      const uint32_t incl = wave_incl_scan(len);  // lanes >= nrows add 0
      // row r covers flattened [start_r, start_r + len_r), start_r = incl_{r-1};
      // lane r holds start_r (lane nrows: the total) and the row's first record
      startv = incl - len;
      rend = incl;
      total = __builtin_amdgcn_readlane(incl, nrows - 1);
      if (W > 1) {
        // the non-empty rows, compacted in row order, with their batches of
        // 64 candidates numbered across the rows
        const uint32_t nb = (len + kWave - 1) / kWave;
        const uint64_t ne = ballot(len > 0);
        const uint32_t binc = wave_incl_scan(nb);
        const int idx = lanes_below(ne);
        if (len > 0) {
          s_rb[idx] = rb;
          s_len[idx] = len;
          s_bstart[idx] = binc - nb;
        }
        if (lane == 0) {
          const uint32_t nrc = static_cast<uint32_t>(__popcll(ne));
          const uint32_t totb = __builtin_amdgcn_readlane(binc, nrows - 1);
          s_bstart[nrc] = totb;
          s_count[0] = nrc;
          s_count[1] = totb;
          s_count[2] = total;
        }
      }
    }
    uint32_t nrc = 0, totb = 0, rbv = 0, lenv = 0, bsv = 0;
    if (W > 1) {
      __syncthreads();
      nrc = __builtin_amdgcn_readfirstlane(s_count[0]);
      totb = __builtin_amdgcn_readfirstlane(s_count[1]);
      total = __builtin_amdgcn_readfirstlane(s_count[2]);
      // lane r holds compacted row r (lane nrc: the batch total)
      if (lane < static_cast<int>(nrc)) {
        rbv = s_rb[lane];
        lenv = s_len[lane];
      }
      if (lane <= static_cast<int>(nrc))
        bsv = s_bstart[lane];
    }
    if (wsub == 0)
      cand += total;
    R acc[D];
    for (int k = 0; k < D; ++k)
      acc[k] = 0;
    R wl = 0;
    uint32_t in32 = 0;
    // Candidates of one batch are consecutive in a row of cells, so whole
    // batches tend to fall outside the support together: the branch lets
    // the wave skip the weight.  The sentinel record is never inside.
    auto test = [&](const typename Vec<R>::V4 &p, const typename Vec<R>::V2 &q) {
      R d[D];
      d[0] = p.x - m[0];
      d[1] = p.y - m[1];
      d[2] = p.z - m[2];
      d[3] = p.w - m[3];
      if (D == 6) {
        d[4 % D] = q.x - m[4 % D];
        d[5 % D] = q.y - m[5 % D];
      }
      R dd = d[0] * d[0];
      for (int k = 1; k < D; ++k)
        dd = ms_fma(d[k], d[k], dd);
      // in-support visits: the wave's count, kept in a scalar register
      const bool inb = dd < r2;
      in32 += static_cast<uint32_t>(__popcll(ballot(inb)));
      if (inb) {
        const R w = ms_gauss<R>(dd, gscale) - floor_w;
        wl += w;
        for (int k = 0; k < D; ++k)
          acc[k] = ms_fma(w, d[k], acc[k]);
      }
    };
    if constexpr (W == 1) {
      // Chunks of the flattened range: the map (LDS) sends candidate c to
      // its record index, written row by row, then the wave walks the chunk
      // in 64-candidate batches, batch k + 1 loading while batch k is tested.
      // Loads are unconditional: lanes past the chunk's end read the map's
      // last slot, the sentinel record, which lies outside every support.
      constexpr uint32_t kChunk = kMapCap - 1;  // the last slot holds the sentinel
      for (uint32_t cbase = 0; cbase < total; cbase += kChunk) {
        const uint32_t cend = total - cbase < kChunk ? total : cbase + kChunk;
        // With short rows (at most kShort candidates a row on average:
        // small meshes), each lane writes the first kShort slots of its own
        // row and the wave fills the rest of the longer rows; else the wave
        // fills the rows one by one.
        constexpr uint32_t kShort = 8;
        if (total <= kShort * static_cast<uint32_t>(nrows)) {
          const uint32_t lo = startv > cbase ? startv : cbase, hi = rend < cend ? rend : cend;
          for (uint32_t e = 0; e < kShort; ++e)
            if (lo + e < hi)
              map[lo + e - cbase] = rb + (lo + e - startv);
          uint64_t longer = ballot(hi > lo + kShort);
          while (longer) {
            const int rr = __builtin_ctzll(longer);
            longer &= longer - 1;
            const uint32_t rs = __builtin_amdgcn_readlane(startv, rr);
            const uint32_t ra = __builtin_amdgcn_readlane(lo, rr) + kShort;
            const uint32_t rz = __builtin_amdgcn_readlane(hi, rr);
            const uint32_t b = __builtin_amdgcn_readlane(rb, rr);
            for (uint32_t c = ra + lane; c < rz; c += kWave)
              map[c - cbase] = b + (c - rs);
          }
        } else {
          for (int rr = 0; rr < nrows; ++rr) {
            const uint32_t rs = __builtin_amdgcn_readlane(startv, rr);
            const uint32_t re = __builtin_amdgcn_readlane(startv, rr + 1);
            const uint32_t lo = rs > cbase ? rs : cbase, hi = re < cend ? re : cend;
            const uint32_t b = __builtin_amdgcn_readlane(rb, rr);
            for (uint32_t c = lo + lane; c < hi; c += kWave)
              map[c - cbase] = b + (c - rs);
          }
        }
        const uint32_t send = cend - cbase;  // the sentinel's slot
        if (lane == 0)
          map[send] = sentinel;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        // three stages per batch: map slot (LDS), record (global), test; the
        // map read runs two batches ahead and the record load one batch ahead
        auto slot = [&](uint32_t c) -> uint32_t {
          return map[c - cbase < send ? c - cbase : send];
        };
        auto load = [&](uint32_t j, typename Vec<R>::V4 &p, typename Vec<R>::V2 &q) {
          p = p4[j];
          if (D == 6)
            q = q2[j];
        };
        uint32_t c0 = cbase;
        if (c0 < cend) {
          typename Vec<R>::V4 pa, pb;
          typename Vec<R>::V2 qa{}, qb{};
          uint32_t ja = slot(c0 + lane), jb = slot(c0 + kWave + lane);
          load(ja, pa, qa);
          for (; c0 < cend; c0 += 2 * kWave) {
            // the scheduling barriers keep each prefetch ahead of the tests
            // (a load sunk into the iteration that reads it waits in full)
            load(jb, pb, qb);
            ja = slot(c0 + 2 * kWave + lane);
            __builtin_amdgcn_sched_barrier(0);
            test(pa, qa);
            load(ja, pa, qa);
            jb = slot(c0 + 3 * kWave + lane);
            __builtin_amdgcn_sched_barrier(0);
            test(pb, qb);
          }
        }
        __builtin_amdgcn_wave_barrier();  // the map is rewritten by the next chunk
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      }
    } else {
      // Row-direct walk: batch k of the numbering above is 64 consecutive
      // records of one row.  Wave w takes the contiguous batches [k0, k1) of
      // a W-way split, so its row cursor advances over about a W-th of the
      // rows; it starts at the row holding k0 (the rows whose first batch is
      // <= k0, counted by one ballot).  Lanes past the row's end read the
      // sentinel record, which lies outside every support.  The next batch's
      // records are in flight while the current batch is tested.
      const uint32_t k0 = static_cast<uint32_t>(
                         (static_cast<uint64_t>(totb) * static_cast<uint32_t>(wsub)) / W),
                     k1 = static_cast<uint32_t>(
                         (static_cast<uint64_t>(totb) * static_cast<uint32_t>(wsub + 1)) / W);
      uint32_t r = 0, bs_cur = 0, bs_next = 0, rb_r = 0, len_r = 0;
      if (k0 < k1) {
        const uint64_t below = ballot(lane < static_cast<int>(nrc) && bsv <= k0);
        r = static_cast<uint32_t>(__popcll(below)) - 1;
        bs_cur = __builtin_amdgcn_readlane(bsv, r);
        bs_next = __builtin_amdgcn_readlane(bsv, r + 1);
        rb_r = __builtin_amdgcn_readlane(rbv, r);
        len_r = __builtin_amdgcn_readlane(lenv, r);
      }
      // batch k -> each lane's record, as a byte offset into p4 (the q2
      // record sits at half that offset): one add of the lane's own offset
      // to the batch's scalar base, one compare and one select per batch,
      // and the loads take the array as a scalar base
      constexpr uint32_t kRec = sizeof(typename Vec<R>::V4);
      const uint32_t lane_b = static_cast<uint32_t>(lane) * kRec, sent_b = sentinel * kRec;
      auto rec_of = [&](uint32_t k) -> uint32_t {
        if (k >= k1)
          return sent_b;
        while (k >= bs_next) {
          ++r;
          bs_cur = bs_next;
          bs_next = __builtin_amdgcn_readlane(bsv, r + 1);
          rb_r = __builtin_amdgcn_readlane(rbv, r);
          len_r = __builtin_amdgcn_readlane(lenv, r);
        }
        const uint32_t off = (k - bs_cur) * kWave;
        const uint32_t base = (rb_r + off) * kRec;
        return static_cast<uint32_t>(lane) < len_r - off ? base + lane_b : sent_b;
      };
      auto load = [&](uint32_t b, typename Vec<R>::V4 &p, typename Vec<R>::V2 &q) {
        p = *reinterpret_cast<const typename Vec<R>::V4 *>(reinterpret_cast<const char *>(p4) + b);
        if (D == 6)
          q = *reinterpret_cast<const typename Vec<R>::V2 *>(reinterpret_cast<const char *>(q2) +
                                                             (b >> 1));
      };
      uint32_t k = k0;
      if (k < k1) {
        typename Vec<R>::V4 pa, pb;
        typename Vec<R>::V2 qa{}, qb{};
        load(rec_of(k), pa, qa);
        uint32_t jb = rec_of(k + 1);
        for (; k < k1; k += 2) {
          load(jb, pb, qb);
          const uint32_t ja = rec_of(k + 2);
          __builtin_amdgcn_sched_barrier(0);
          test(pa, qa);
          load(ja, pa, qa);
          jb = rec_of(k + 3);
          __builtin_amdgcn_sched_barrier(0);
          test(pb, qb);
        }
      }
    }
    inside += in32;
    ++done;
    bool moved = false;
    if (W == 1) {
      wl = wave_total(wl);
      for (int k = 0; k < D; ++k)
        acc[k] = wave_total(acc[k]);
      wsum = wl;
      if (wl > R(0)) {
        const R inv = R(1) / wl;
        for (int k = 0; k < D; ++k) {
          const R nk = m[k] + acc[k] * inv;
          moved = moved || nk != m[k];
          m[k] = nk;
        }
      }
    } else {
      // lane-wise partials through LDS; component c (the D sums, then the
      // weight) is summed over the waves in a fixed order by wave c mod W,
      // then every wave takes the D + 1 totals and makes the same update
      for (int k = 0; k < D; ++k)
        red[k * kThreads + threadIdx.x] = acc[k];
      red[D * kThreads + threadIdx.x] = wl;
      __syncthreads();
      for (int c = wave; c <= D; c += W) {
        const R *row = red + c * kThreads + lane;
        const R v = wave_total((row[0] + row[kWave]) + (row[2 * kWave] + row[3 * kWave]));
        if (lane == 0)
          s_mean[c] = v;
      }
      __syncthreads();
      R v[D + 1];
      for (int k = 0; k <= D; ++k)
        v[k] = s_mean[k];
      wl = v[D];
      wsum = wl;
      if (wl > R(0)) {
        const R inv = R(1) / wl;
        for (int k = 0; k < D; ++k) {
          const R nk = m[k] + v[k] * inv;
          moved = moved || nk != m[k];
          m[k] = nk;
        }
      }
      // no barrier here: the totals are rewritten only after the next
      // iteration's two barriers, which every wave reaches after reading them
    }
    if (!moved)
      break;  // uniform across the seed's waves: they hold the same mean
  }
  if (W > 1) {
    __shared__ unsigned long long ins[kWavesPerBlock];
    if (lane == 0)
      ins[wave] = inside;
    __syncthreads();
    inside = ins[0] + ins[1] + ins[2] + ins[3];
  }
  if (lane == 0 && wsub == 0) {
    for (int k = 0; k < D; ++k)
      out[g.perm[k] * ns + s] = m[k];
    weight[s] = wsum;
    stats[s] = cand;  // per-seed slots: no contended atomics
    stats[ns + s] = inside;
    stats[2 * ns + s] = static_cast<unsigned long long>(done);
  }
}

// Sum over each G-lane group (G = 16: a DPP row; G = 32: two rows), the
// same bits in every lane of the group: exchanges between lanes i and i ^ 1,
// i ^ 2, the quads of a half row (uniform by then), the halves of a row and
// (G = 32) the rows of a pair.  IEEE addition is commutative, so both lanes
// of each exchange add the same two values.
template <int G, class R> __device__ inline R group_total(R v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror
  if (G == 32)
    v += __shfl_xor(v, 16, kWave);
  return v;
}

template <int G> __device__ inline uint32_t group_count(uint32_t v) {
  v += dpp_u32<0xB1>(0u, v);
  v += dpp_u32<0x4E>(0u, v);
  v += dpp_u32<0x141>(0u, v);
  v += dpp_u32<0x140>(0u, v);
  if (G == 32)
    v += static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), 16, kWave));
  return v;
}

// Mean-shift queries with G lanes per seed (64 / G seeds per wave, 256 / G
// per block), for many seeds with few candidates each (one-wave seeds would
// spend most of their issue slots on the per-iteration work: the mean's cell,
// the row ranges, the reductions, the update).  Grids of reach 1: 9 rows, one
// per lane of the group's first 16.  The rows' records are numbered one after
// the other (row order) and written, as byte offsets, into the group's map in
// LDS; the group then walks the map in batches of G entries, so a batch spans
// rows and no lane idles at a short row's end.  Rows are kept while the mean
// stays in the (y, z) cell they were found for and within kReuse * cutoff * h
// of the mean they were cut around (the anchor): they are cut to the support
// ball widened by that distance, which holds every point of the support of
// any mean in that range, and the map is kept too.  Loops are wave-uniform
// (the most any group of the wave needs); lanes past their group's end read
// the sentinel record.  Group sums come from group_total, so every lane of a
// group holds the same new mean; a seed whose update leaves it unchanged
// stops, and the wave leaves once all its seeds have.
constexpr int kSmallRows = (2 * kNarrowReach + 1) * (2 * kNarrowReach + 1);
constexpr uint32_t kSmallMapBlock = 4096;  // LDS map entries per block (16 KiB)
constexpr double kReuse = 0.1;
// kSeg: seeds of several segments (meshes) with a grid each; otherwise every
// seed shares grid 0, held in scalar registers.
template <class R, int D, int G, bool kSeg>
__global__ void __launch_bounds__(kThreads, sizeof(R) == 4 ? 5 : 1)  // FP32: <= 96 VGPRs
    k_ms_query_small(const typename Vec<R>::V4 *p4, const typename Vec<R>::V2 *q2,
                     const uint64_t *offsets, const Grid *grids, const uint64_t *cellbase,
                     const uint32_t *seed_mesh, const uint64_t *ns_dev,
                     const uint32_t *seed_pos, const R *seeds, const uint32_t *order,
                     uint64_t ns, unsigned logical_blocks, uint32_t sentinel, int iterations,
                     R gscale, R r2, R floor_w, R *out, R *weight,
                     unsigned long long *stats) {
  static_assert(G == 16 || G == 32, "16 or 32 lanes per seed");
  static_assert(kSmallRows <= 16, "the rows fit the first DPP row of a group");
  constexpr int kGroups = kThreads / G;
  constexpr uint32_t kCap = kSmallMapBlock / kGroups;  // map entries per group
  constexpr uint32_t kRec = sizeof(typename Vec<R>::V4);
  __shared__ uint32_t smap[kSmallMapBlock];
  __shared__ uint4 srows[kGroups * kSmallRows];  // (first entry, first byte, length)
  const int lane = threadIdx.x & (kWave - 1);
  const int sl = lane & (G - 1);   // lane in the group
  const int grp = static_cast<int>(threadIdx.x) / G;
  const int wave = threadIdx.x / kWave;
  uint32_t *map = smap + grp * kCap;
  uint4 *rows = srows + grp * kSmallRows;
  const unsigned per = (logical_blocks + kXcds - 1) / kXcds;
  const unsigned logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  if (logical >= logical_blocks)
    return;
  const uint64_t slot = static_cast<uint64_t>(logical) * kGroups +
                        static_cast<uint64_t>(wave) * (kWave / G) + lane / G;
  const uint64_t nseeds = ns_dev ? *ns_dev : ns;
  bool active = slot < nseeds;
  if (!ballot(active))
    return;  // the whole wave
  const uint64_t s = active ? (order ? order[slot] : slot) : 0;
  const uint32_t gm = kSeg && active ? seed_mesh[s] : 0;
  const Grid g = grids[gm];
  const uint64_t cb = cellbase ? cellbase[gm] : 0;
  R m[D];
  for (int k = 0; k < D; ++k)
    m[k] = 0;
  if (active) {
    if (seed_pos) {
      const uint64_t p = seed_pos[s];
      const typename Vec<R>::V4 v = p4[p];
      m[0] = v.x, m[1] = v.y, m[2] = v.z, m[3] = v.w;
      if (D == 6) {
        const typename Vec<R>::V2 u = q2[p];
        m[4 % D] = u.x, m[5 % D] = u.y;
      }
    } else {
      for (int k = 0; k < D; ++k)
        m[k] = seeds[g.perm[k] * ns + s];
    }
  }
  R wsum = 0;
  unsigned long long cand = 0;
  uint32_t inside = 0;
  int done = 0;
  // rows cut to the support ball widened by the reuse distance (and a
  // relative margin over the FP64 test of that distance below)
  const double reuse = kReuse * sqrt(static_cast<double>(r2));
  RowCut rc = row_cut(g, static_cast<double>(r2));
  {
    const double rl = sqrt(rc.rlim2) + reuse * (1.0 + 1e-6);
    rc.rlim2 = rl * rl;
  }
  const double reuse2 = reuse * reuse;
  const uint32_t sent_b = sentinel * kRec;
  R anc[3] = {R(0), R(0), R(0)};  // the anchor's grid coordinates
  int acy = -1, acz = -1;
  uint32_t rb = 0, len = 0;       // lanes 0..8 of a group: its row's first record, length
  uint32_t gtot = 0, chunks = 0;  // the group's entries; map chunks (wave-uniform)
  bool mapped = false;            // the map holds the rows' single chunk
  for (int it = 0; it < iterations; ++it) {
    const double mx = static_cast<double>(m[0]), my = static_cast<double>(m[1]),
                 mz = static_cast<double>(m[2]);
    const int cy = cell_of(my, g.lo[1], g.inv, g.dims[1]);
    const int cz = cell_of(mz, g.lo[2], g.inv, g.dims[2]);
    const double ex = mx - static_cast<double>(anc[0]), ey = my - static_cast<double>(anc[1]),
                 ez = mz - static_cast<double>(anc[2]);
    const bool need = active && (cy != acy || cz != acz || ex * ex + ey * ey + ez * ez > reuse2);
    if (ballot(need)) {
      // the groups that need it cut their rows around their means (cut_row:
      // lanes 0..8 of the group, one row each; a seed's rows depend on its
      // own path only, so its sums do not depend on the other seeds of the
      // wave), then every group numbers its rows' records by a prefix over
      // its first DPP row and lists them in LDS
      if (need) {
        anc[0] = m[0], anc[1] = m[1], anc[2] = m[2];
        acy = cy, acz = cz;
        rb = 0, len = 0;
        if (sl < kSmallRows)
          cut_row(g, rc, cb, offsets, mx, my, mz, cy, cz, sl, rb, len);
      }
      if (!active)
        len = 0;
      uint32_t incl = len;
      incl += dpp_u32<0x111>(0u, incl);  // row_shr 1, 2, 4, 8
      incl += dpp_u32<0x112>(0u, incl);
      incl += dpp_u32<0x114>(0u, incl);
      incl += dpp_u32<0x118>(0u, incl);
      if (sl < kSmallRows)
        rows[sl] = make_uint4(incl - len, rb * kRec, len, 0u);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      const uint4 last = rows[kSmallRows - 1];
      gtot = last.x + last.z;
      uint32_t mc = 0;
      for (int q = 0; q < kWave / G; ++q)
        mc = max(mc, static_cast<uint32_t>(__builtin_amdgcn_readlane(gtot, q * G)));
      chunks = (mc + kCap - 1) / kCap;
      mapped = false;
    }
    const uint32_t mine = active ? gtot : 0u;
    cand += mine;
    R acc[D];
    for (int k = 0; k < D; ++k)
      acc[k] = 0;
    R wl = 0;
    auto test = [&](const typename Vec<R>::V4 &p, const typename Vec<R>::V2 &q) {
      R d[D];
      d[0] = p.x - m[0];
      d[1] = p.y - m[1];
      d[2] = p.z - m[2];
      d[3] = p.w - m[3];
      if (D == 6) {
        d[4 % D] = q.x - m[4 % D];
        d[5 % D] = q.y - m[5 % D];
      }
      R dd = d[0] * d[0];
      for (int k = 1; k < D; ++k)
        dd = ms_fma(d[k], d[k], dd);
      if (dd < r2) {
        const R w = ms_gauss<R>(dd, gscale) - floor_w;
        wl += w;
        for (int k = 0; k < D; ++k)
          acc[k] = ms_fma(w, d[k], acc[k]);
        ++inside;
      }
    };
    auto load = [&](uint32_t b, typename Vec<R>::V4 &p, typename Vec<R>::V2 &q) {
      p = *reinterpret_cast<const typename Vec<R>::V4 *>(reinterpret_cast<const char *>(p4) + b);
      if (D == 6)
        q = *reinterpret_cast<const typename Vec<R>::V2 *>(reinterpret_cast<const char *>(q2) +
                                                           (b >> 1));
    };
    for (uint32_t ch = 0; ch < chunks; ++ch) {
      const uint32_t c0 = ch * kCap;
      // this group's entries in the chunk (none once its seed has stopped)
      const uint32_t glen = mine > c0 ? min(mine - c0, kCap) : 0u;
      if (!mapped) {
        // entry f of row r (first entry x, first byte y, length z) is the
        // record at byte y + (f - x) * kRec
#pragma unroll 1
        for (int r = 0; r < kSmallRows; ++r) {
          const uint4 t = rows[r];
          const uint32_t lo = max(t.x, c0), hi = min(t.x + t.z, c0 + kCap);
          for (uint32_t f = lo + sl; f < hi; f += G)
            map[f - c0] = t.y + (f - t.x) * kRec;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      }
      uint32_t nb = 0;
      for (int q = 0; q < kWave / G; ++q)
        nb = max(nb, static_cast<uint32_t>(__builtin_amdgcn_readlane(glen, q * G)));
      nb = (nb + G - 1) / G;
      // three stages per batch: map entry (LDS), record (global), test; the
      // map read runs two batches ahead and the record load one batch ahead
      auto entry = [&](uint32_t t) -> uint32_t {
        const uint32_t i = t * G + static_cast<uint32_t>(sl);
        const uint32_t v = map[i < glen ? i : 0u];
        return i < glen ? v : sent_b;
      };
      if (nb) {
        typename Vec<R>::V4 pa, pb;
        typename Vec<R>::V2 qa{}, qb{};
        uint32_t ja = entry(0), jb = entry(1);
        load(ja, pa, qa);
        for (uint32_t t = 0; t < nb; t += 2) {
          load(jb, pb, qb);
          ja = entry(t + 2);
          __builtin_amdgcn_sched_barrier(0);
          test(pa, qa);
          load(ja, pa, qa);
          jb = entry(t + 3);
          __builtin_amdgcn_sched_barrier(0);
          test(pb, qb);
        }
      }
      if (chunks > 1) {  // the next chunk rewrites the map
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      }
    }
    mapped = chunks == 1;
    wl = group_total<G>(wl);
    for (int k = 0; k < D; ++k)
      acc[k] = group_total<G>(acc[k]);
    bool moved = false;
    if (active) {
      wsum = wl;
      ++done;
      if (wl > R(0)) {
        const R inv = R(1) / wl;
        for (int k = 0; k < D; ++k) {
          const R nk = m[k] + acc[k] * inv;
          moved = moved || nk != m[k];
          m[k] = nk;
        }
      }
    }
    active = active && moved;  // group-uniform: every lane holds the same sums
    if (!ballot(active))
      break;
  }
  const uint32_t ins = group_count<G>(inside);
  if (sl == 0 && slot < nseeds) {
    for (int k = 0; k < D; ++k)
      out[g.perm[k] * ns + s] = m[k];
    weight[s] = wsum;
    stats[s] = cand;
    stats[ns + s] = ins;
    stats[2 * ns + s] = static_cast<unsigned long long>(done);
  }
}

__global__ void k_sum_u64(const unsigned long long *a, uint64_t n, unsigned long long *out) {
  __shared__ unsigned long long sh[kThreads];
  unsigned long long x = 0;
  for (uint64_t i = threadIdx.x; i < n; i += kThreads)
    x += a[i];
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *out = sh[0];
}

}  // namespace

// Control block of the session's mean-shift calls (slot kSlotMs1, after the
// per-block box partials).
static MsCtl *ms_ctl(Session *s) {
  const size_t part = kPart * kBboxBlocks * sizeof(double);
  char *base = static_cast<char *>(scratch(s, kSlotMs1, part + sizeof(MsCtl)));
  return reinterpret_cast<MsCtl *>(base + part);
}

void ms_begin(Session *s, bool fork) {
  s->ms_held.clear();
  zero_device(s, ms_ctl(s), sizeof(MsCtl));
  if (!fork)
    return;
  // the fork point of a side run: after the call's inputs and the cleared
  // control block, before any run's work on the session stream
  if (!s->side_stream) {
    MVS_HIP(hipStreamCreateWithFlags(&s->side_stream, hipStreamNonBlocking));
    MVS_HIP(hipEventCreateWithFlags(&s->fork_event, hipEventDisableTiming));
    MVS_HIP(hipEventCreateWithFlags(&s->join_event, hipEventDisableTiming));
  }
  MVS_HIP(hipEventRecord(s->fork_event, s->stream));
}

void ms_release(Session *s) {
  std::vector<Session::MsHeld> held;
  held.swap(s->ms_held);
  const size_t k = held.size();
  if (k > 1) {
    for (size_t i = 0; i < k; ++i) {
      if (!s->setup_events[i])
        MVS_HIP(hipEventCreateWithFlags(&s->setup_events[i], hipEventDisableTiming));
      MVS_HIP(hipEventRecord(s->setup_events[i], held[i].stream));
    }
    for (size_t i = 0; i < k; ++i)
      for (size_t j = 0; j < k; ++j)
        if (j != i && held[j].stream != held[i].stream && held[i].gate)
          MVS_HIP(hipStreamWaitEvent(held[i].stream, s->setup_events[j], 0));
  }
  const hipStream_t main = s->stream;
  const int bank = s->bank;
  struct Restore {
    Session *s;
    hipStream_t main;
    int bank;
    ~Restore() {
      s->stream = main;
      s->bank = bank;
    }
  } restore{s, main, bank};
  for (auto &q : held) {
    s->stream = q.stream;
    s->bank = q.bank;
    q.launch();
  }
}

unsigned *ms_bad_flag(Session *s) { return &ms_ctl(s)->bad; }

void ms_finish(Session *s, unsigned long long sums[kMsSums]) {
  ms_release(s);
  if (s->side_pending) {  // the side stream's query and sums come first
    MVS_HIP(hipStreamWaitEvent(s->stream, s->join_event, 0));
    s->side_pending = false;
  }
  // the visit sums and the flag only (a captured step compares what it reads)
  struct {
    unsigned long long sums[kMsSums];
    unsigned bad;
  } h;
  static_assert(offsetof(MsCtl, bad) ==
                    offsetof(MsCtl, sums) + kMsSums * sizeof(unsigned long long),
                "sums and flag are adjacent");
  read_back(s, &h, reinterpret_cast<const char *>(ms_ctl(s)) + offsetof(MsCtl, sums),
            kMsSums * sizeof(unsigned long long) + sizeof(unsigned));
  if (h.bad)
    raise(MVS_EINVAL, "non-finite mean-shift points");
  if (sums)
    std::memcpy(sums, h.sums, sizeof h.sums);
}

// Builds the grid(s) on the device and runs every seed, with no host
// synchronisation; visit sums go to ctl->sums[run_slot], [run_slot + 1].
// Seeds: explicit positions (`seeds`, D x ns SoA), or, when seed_ids is
// given, ns points drawn from the sorted points (salt), whose indices are
// written to seed_ids.  With segments (batches of meshes: seg->meshes
// grids, drawn seeds only) ns is the seed capacity and seg->seed_off holds
// the per-mesh seed counts' starts on the device.
template <class R>
static void ms_batched(Session *s, int D, uint64_t n, const R *pts, uint64_t ns, const R *seeds,
                       int32_t *seed_ids, uint64_t salt, int iterations, double h, double cutoff,
                       R *out, R *weight, int run_slot, const MsSegs *seg, MsRun run) {
  using V4 = typename Vec<R>::V4;
  using V2 = typename Vec<R>::V2;
  // slots read by the query (the set-up's own scratch is shared)
  const int slot_rec = run.set ? kSlotMsB0 : kSlotMs0, slot_off = run.set ? kSlotMsB5 : kSlotMs5,
            slot_seed = run.set ? kSlotMsB6 : kSlotMs6, slot_stats = run.set ? kSlotMsB7 : kSlotMs7;
  if (D != 4 && D != 6)
    raise(MVS_EINVAL, "batched mean-shift supports 4-D (reflection) and 6-D (rigid) points");
  if (!(h > 0) || !(cutoff > 0) || !std::isfinite(h * cutoff) || iterations < 0)
    raise(MVS_EINVAL, "invalid mean-shift parameters");
  // the queries address records by 32-bit byte offsets (the sentinel at [n])
  if (n > static_cast<uint64_t>(UINT32_MAX) / sizeof(V4) - 1 ||
      ns > static_cast<uint64_t>(UINT32_MAX))
    raise(MVS_EOVERFLOW, "too many points or seeds");
  if (seed_ids && !seg && ns > n)
    raise(MVS_EINVAL, "more seed votes than votes");
  if (seg && !seed_ids)
    raise(MVS_EINVAL, "segmented mean-shift draws its seeds");
  if (ns == 0)
    return;
  // a side run forks from the session stream here: its set-up (in the bank-1
  // scratch, sort and scan slots) and its query run beside what the session
  // stream does next; ms_finish joins it
  const hipStream_t main = s->stream;
  const int main_bank = s->bank;
  if (run.side) {  // from the fork point recorded by ms_begin
    if (!s->side_stream)
      raise(MVS_EINVAL, "a side mean-shift run needs ms_begin(s, true)");
    MVS_HIP(hipStreamWaitEvent(s->side_stream, s->fork_event, 0));
    s->stream = s->side_stream;
  }
  s->bank = run.set;
  struct Restore {
    Session *s;
    hipStream_t main;
    int bank;
    ~Restore() {
      s->stream = main;
      s->bank = bank;
    }
  } restore{s, main, main_bank};
  const bool bank1 = run.set != 0;
  const uint32_t meshes = seg ? seg->meshes : 1;
  const uint64_t *seg_off = seg ? seg->seg_off : nullptr;
  MsCtl *ctl = ms_ctl(s);
  // per-seed statistics [0, ns) candidates, [ns, 2ns) in-support visits,
  // [2ns, 3ns) updates computed (a seed stops at a fixed point)
  unsigned long long *stats = scratch_as<unsigned long long>(s, slot_stats, 3 * ns);
  if (seg)  // seeds past the drawn count leave their slots untouched
    zero_device(s, stats, 3 * ns * sizeof(unsigned long long));
  const uint64_t np = n ? n : 1;
  // cell-sorted records, then the sentinel record at [n]
  char *sorted =
      static_cast<char *>(scratch(s, slot_rec, (np + 1) * (sizeof(V4) + sizeof(V2))));
  V4 *p4 = reinterpret_cast<V4 *>(sorted);
  V2 *q2 = reinterpret_cast<V2 *>(sorted + (np + 1) * sizeof(V4));
  // cells: at most clamp(2 n, 1024, kMaxCells) for one point set (the bound
  // k_grid_setup lays its grid under), the caller's bound for segments
  const uint64_t one_cap = n < kMaxCells / 2 ? std::max<uint64_t>(2 * n, 1024) : kMaxCells;
  const uint64_t cells_cap = seg ? std::max<uint64_t>(seg->cells_bound, 1) : one_cap;
  uint64_t *off = scratch_as<uint64_t>(s, slot_off, cells_cap + 1);
  Grid *grids = seg ? scratch_as<Grid>(s, kSlotBatchMs0, meshes)
                    : (run.set ? &ctl->grid2 : &ctl->grid);
  uint64_t *cellbase = nullptr;
  const uint64_t *total_cells = reinterpret_cast<const uint64_t *>(
      reinterpret_cast<const char *>(grids) + offsetof(Grid, ncells));
  // boxes and moments: per-block partials of the one point set, or one
  // block per segment
  unsigned nb = 0;
  double *partial = nullptr;
  if (!seg) {
    // bank 0 shares its slot with the control block, which follows the partials
    partial = scratch_as<double>(s, bank1 ? kSlotMsBankPart : kSlotMs1, kPart * kBboxBlocks);
    if (n) {
      const R *fa = static_cast<const R *>(run.from_a), *fb = static_cast<const R *>(run.from_b);
      // points from votes: a round of kPointsPerRound<R> points per thread
      nb = blocks_for(n, fa ? kThreads * kPointsPerRound<R> : kThreads);
      if (nb > kBboxBlocks)
        nb = kBboxBlocks;
      R *wpts = const_cast<R *>(pts);  // the caller's point buffer, written here
      if (fa && D == 6)
        MVS_LAUNCH(s, "ms_points", (k_points_moments<R, 6>), nb, kThreads, 0, fa, fb, n, n,
                   wpts, partial);
      else if (fa)
        MVS_LAUNCH(s, "ms_points", (k_points_moments<R, 4>), nb, kThreads, 0, fa, fb,
                   run.from_stride, n, wpts, partial);
      else if (D == 6)
        MVS_LAUNCH(s, "ms_bbox", (k_bbox<R, 6>), nb, kThreads, 0, pts, n, partial);
      else
        MVS_LAUNCH(s, "ms_bbox", (k_bbox<R, 4>), nb, kThreads, 0, pts, n, partial);
    }
  } else {
    partial = scratch_as<double>(s, kSlotBatchMs1, static_cast<size_t>(kPart) * meshes * kSegBlocks);
    nb = kSegBlocks;
    if (D == 6)
      MVS_LAUNCH(s, "ms_bbox", (k_seg_moments<R, 6>), meshes * kSegBlocks, kThreads, 0, pts, n,
                 seg_off, partial);
    else
      MVS_LAUNCH(s, "ms_bbox", (k_seg_moments<R, 4>), meshes * kSegBlocks, kThreads, 0, pts, n,
                 seg_off, partial);
  }
  uint64_t *ncell = seg ? scratch_as<uint64_t>(s, kSlotBatchMs2, meshes + 1) : nullptr;
  // few seeds: four waves per seed, so the launch still fills the chip.  A
  // seed with a block of its own walks finer cells (reach 3, 49 rows); a
  // narrow seed (16 to 64 lanes) pays for every row, so it walks 9 rows of
  // cells cutoff*h wide, a quarter of that along the rows (reach 1, xdiv 4).
  // Both cut each row to the support ball.
  const bool wide = ns < s->tune.wide_seeds;
  const int max_reach = wide ? kMaxReach : kNarrowReach;
  const int xdiv = wide ? kMaxReach : kNarrowXdiv;
  if (D == 6)
    MVS_LAUNCH(s, "ms_grid", k_grid_setup<6>, meshes, kThreads, 0, partial, nb, seg_off, n,
               cutoff * h, max_reach, xdiv, grids, ncell, &ctl->bad);
  else
    MVS_LAUNCH(s, "ms_grid", k_grid_setup<4>, meshes, kThreads, 0, partial, nb, seg_off, n,
               cutoff * h, max_reach, xdiv, grids, ncell, &ctl->bad);
  int key_bits = kCellBits;
  if (seg) {  // cells of mesh m follow those of meshes < m
    cellbase = scratch_as<uint64_t>(s, kSlotBatchMs3, meshes + 1);
    exclusive_scan_u64(s, ncell, cellbase, meshes);
    total_cells = cellbase + meshes;
    key_bits = 1;  // room for the key cells_cap, past every cell
    while (key_bits < 63 && (1ULL << key_bits) <= cells_cap)
      ++key_bits;
  }
  uint32_t *sorted_ids = nullptr;
  const uint32_t *seed_drawn = nullptr;  // seeds drawn in the launch after the sort
  if (n) {
    uint64_t *keys = scratch_as<uint64_t>(s, bank1 ? kSlotMsBankKeys : kSlotMs2, 2 * n);
    uint32_t *vals = scratch_as<uint32_t>(s, bank1 ? kSlotMsBankVals : kSlotMs3, 2 * n);
    R *aos = scratch_as<R>(s, bank1 ? kSlotMsBankAos : kSlotMs8, kAos * n);
    if (D == 6)
      MVS_LAUNCH(s, "ms_cell_keys", (k_cell_keys<R, 6>), blocks_for(n, kThreads), kThreads, 0,
                 pts, n, grids, seg_off, meshes, cellbase, cells_cap, keys, vals, aos);
    else
      MVS_LAUNCH(s, "ms_cell_keys", (k_cell_keys<R, 4>), blocks_for(n, kThreads), kThreads, 0,
                 pts, n, grids, seg_off, meshes, cellbase, cells_cap, keys, vals, aos);
    radix_sort_pairs_u64(s, keys, vals, keys + n, vals + n, n, key_bits);
    sorted_ids = vals + n;
    // offsets, gathered records and (drawn seeds) the seeds' positions
    SortedArgs A{keys + n, total_cells, off, seg_off, seg ? seg->seed_off : nullptr, meshes,
                 n, ns, salt, sorted_ids, nullptr, seed_ids, seg ? seg->seed_mesh : nullptr,
                 grids};
    const unsigned nb_off = std::min<unsigned>(blocks_for(n + 1, kThreads), 16384u),
                   nb_gather = blocks_for(n, kThreads);
    unsigned nb_pick = 0;
    if (seed_ids) {
      A.pos = scratch_as<uint32_t>(s, slot_seed, ns);
      nb_pick = blocks_for(ns, kThreads);
      seed_drawn = A.pos;
    }
    if (D == 6)
      MVS_LAUNCH(s, "ms_after_sort", (k_after_sort<R, 6>), nb_off + nb_gather + nb_pick, kThreads,
                 0, nb_off, nb_gather, A, aos, p4, q2);
    else
      MVS_LAUNCH(s, "ms_after_sort", (k_after_sort<R, 4>), nb_off + nb_gather + nb_pick, kThreads,
                 0, nb_off, nb_gather, A, aos, p4, q2);
  } else {
    MVS_LAUNCH(s, "ms_cell_offsets", k_cell_offsets, 1, kThreads, 0, nullptr, 0, total_cells,
               off);
  }
  // seeds in cell order
  const uint32_t *order = nullptr, *seed_pos = nullptr;
  if (seed_drawn) {
    seed_pos = seed_drawn;
  } else if (seed_ids) {
    uint32_t *pos = scratch_as<uint32_t>(s, slot_seed, ns);
    MVS_LAUNCH(s, "ms_seeds", k_pick_sorted, blocks_for(ns, kThreads), kThreads, 0, seg_off,
               seg ? seg->seed_off : nullptr, meshes, n, ns, salt, sorted_ids, pos, seed_ids,
               seg ? seg->seed_mesh : nullptr);
    seed_pos = pos;
  } else if (n && ns > 1) {
    uint64_t *skeys = scratch_as<uint64_t>(s, bank1 ? kSlotMsBankSeedKeys : kSlotMs4, 2 * ns);
    uint32_t *svals = scratch_as<uint32_t>(s, slot_seed, 2 * ns);
    MVS_LAUNCH(s, "ms_seed_keys", k_seed_keys<R>, blocks_for(ns, kThreads), kThreads, 0, seeds,
               ns, grids, skeys, svals);
    radix_sort_pairs_u64(s, skeys, svals, skeys + ns, svals + ns, ns, kCellBits);
    order = svals + ns;
  }
  // the Gaussian's scale, folded as ms_gauss takes it
  const R gscale = static_cast<R>(sizeof(R) == 4 ? -1.4426950408889634 / (2.0 * h * h)
                                                 : -1.0 / (2.0 * h * h));
  const R r2 = static_cast<R>(cutoff * h * cutoff * h);
  const R fw = static_cast<R>(std::exp(-0.5 * cutoff * cutoff));
  // narrow queries: 64 (one wave), 32 or 16 lanes per seed
  const unsigned lanes = s->tune.seed_lanes;
  if (lanes != 16 && lanes != 32 && lanes != 64)
    raise(MVS_EINVAL, "lanes per mean-shift seed must be 16, 32 or 64");
  const unsigned logical =
      wide ? static_cast<unsigned>(ns) : blocks_for(ns, kThreads / lanes);
  const unsigned grid = ((logical + kXcds - 1) / kXcds) * kXcds;
  const uint32_t *seed_mesh = seg ? seg->seed_mesh : nullptr;
  const uint64_t *ns_dev = seg ? seg->seed_off + meshes : nullptr;
  const bool side = run.side;
  // the query, its visit sums, and the side run's join (on s->stream)
  auto query = [=]() {
#define MVS_MS_QUERY(DD, WW)                                                                   \
  MVS_LAUNCH(s, "ms_query", (k_ms_query<R, DD, WW>), grid, kThreads, 0, p4, q2, off, grids,   \
             cellbase, seed_mesh, ns_dev, seed_pos, seeds, order, ns, logical,                \
             static_cast<uint32_t>(n), iterations, gscale, r2, fw, out, weight, stats)
#define MVS_MS_SMALL_SEG(DD, GG, SEG)                                                          \
  MVS_LAUNCH(s, "ms_query", (k_ms_query_small<R, DD, GG, SEG>), grid, kThreads, 0, p4, q2,     \
             off, grids, cellbase, seed_mesh, ns_dev, seed_pos, seeds, order, ns, logical,    \
             static_cast<uint32_t>(n), iterations, gscale, r2, fw, out, weight, stats)
#define MVS_MS_SMALL(DD, GG)                                                                   \
  do {                                                                                         \
    if (seed_mesh)                                                                             \
      MVS_MS_SMALL_SEG(DD, GG, true);                                                          \
    else                                                                                       \
      MVS_MS_SMALL_SEG(DD, GG, false);                                                         \
  } while (0)
    if (wide) {
      if (D == 6)
        MVS_MS_QUERY(6, kWavesPerBlock);
      else
        MVS_MS_QUERY(4, kWavesPerBlock);
    } else if (lanes == 64) {
      if (D == 6)
        MVS_MS_QUERY(6, 1);
      else
        MVS_MS_QUERY(4, 1);
    } else if (lanes == 32) {
      if (D == 6)
        MVS_MS_SMALL(6, 32);
      else
        MVS_MS_SMALL(4, 32);
    } else {
      if (D == 6)
        MVS_MS_SMALL(6, 16);
      else
        MVS_MS_SMALL(4, 16);
    }
#undef MVS_MS_SMALL
#undef MVS_MS_SMALL_SEG
#undef MVS_MS_QUERY
    if (run_slot >= 0)
      for (int k = 0; k < 3; ++k)
        MVS_LAUNCH(s, "ms_stats", k_sum_u64, 1, kThreads, 0, stats + k * ns, ns,
                   ctl->sums + 3 * run_slot + k);
    if (side) {
      MVS_HIP(hipEventRecord(s->join_event, s->side_stream));
      s->side_pending = true;
    }
  };
  if (run.hold) {
    if (s->ms_held.size() >= 2)
      raise(MVS_EINVAL, "at most two held mean-shift runs");
    s->ms_held.push_back(Session::MsHeld{s->stream, s->bank, run.gate, query});
  } else {
    query();
  }
}

void ms_batched_f32(Session *s, int D, uint64_t n, const float *pts, uint64_t ns,
                    const float *seeds, int32_t *seed_ids, uint64_t salt, int iterations,
                    double h, double cutoff, float *out, float *weight, int run_slot,
                    const MsSegs *seg, MsRun run) {
  ms_batched<float>(s, D, n, pts, ns, seeds, seed_ids, salt, iterations, h, cutoff, out, weight,
                    run_slot, seg, run);
}
void ms_batched_f64(Session *s, int D, uint64_t n, const double *pts, uint64_t ns,
                    const double *seeds, int32_t *seed_ids, uint64_t salt, int iterations,
                    double h, double cutoff, double *out, double *weight, int run_slot,
                    const MsSegs *seg, MsRun run) {
  ms_batched<double>(s, D, n, pts, ns, seeds, seed_ids, salt, iterations, h, cutoff, out, weight,
                     run_slot, seg, run);
}
void ms_points_device(Session *s, int precision, uint64_t nrigid, const void *q4, const void *t3,
                      uint64_t nrefl, const void *n3, uint64_t n3_stride, const void *d1,
                      void *rigid_pts, void *refl_pts) {
  if (precision == MVS_FP32_GUARDED) {
    if (nrigid)
      MVS_LAUNCH(s, "ms_points", k_points_rigid<float>, blocks_for(nrigid, kThreads), kThreads, 0,
                 static_cast<const float *>(q4), static_cast<const float *>(t3), nrigid,
                 static_cast<float *>(rigid_pts));
    if (nrefl)
      MVS_LAUNCH(s, "ms_points", k_points_refl<float>, blocks_for(nrefl, kThreads), kThreads, 0,
                 static_cast<const float *>(n3), n3_stride, static_cast<const float *>(d1), nrefl,
                 static_cast<float *>(refl_pts));
  } else if (precision == MVS_FP64) {
    if (nrigid)
      MVS_LAUNCH(s, "ms_points", k_points_rigid<double>, blocks_for(nrigid, kThreads), kThreads,
                 0, static_cast<const double *>(q4), static_cast<const double *>(t3), nrigid,
                 static_cast<double *>(rigid_pts));
    if (nrefl)
      MVS_LAUNCH(s, "ms_points", k_points_refl<double>, blocks_for(nrefl, kThreads), kThreads, 0,
                 static_cast<const double *>(n3), n3_stride, static_cast<const double *>(d1), nrefl,
                 static_cast<double *>(refl_pts));
  } else
    raise(MVS_EINVAL, "unknown precision");
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_ms_points(mvs_session *sp, int precision, uint64_t nrigid, const void *q4,
                      const void *t3, uint64_t nrefl, const void *n3, const void *d1,
                      void *rigid_pts, void *refl_pts) {
  return guarded([&] {
    ms_points_device(as_session(sp), precision, nrigid, q4, t3, nrefl, n3, nrefl, d1, rigid_pts,
                     refl_pts);
  });
}

int mvs_dev_ms_batched(mvs_session *sp, int precision, int dim, uint64_t n, const void *pts,
                       uint64_t ns, const void *seeds, int iterations, double h, double cutoff,
                       void *out, void *weight, uint64_t *candidates, uint64_t *in_support,
                       uint64_t *updates) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (precision != MVS_FP32_GUARDED && precision != MVS_FP64)
      raise(MVS_EINVAL, "unknown precision");
    ms_begin(s);
    if (precision == MVS_FP32_GUARDED)
      ms_batched<float>(s, dim, n, static_cast<const float *>(pts), ns,
                        static_cast<const float *>(seeds), nullptr, 0, iterations, h, cutoff,
                        static_cast<float *>(out), static_cast<float *>(weight), 0, nullptr,
                        MsRun{});
    else
      ms_batched<double>(s, dim, n, static_cast<const double *>(pts), ns,
                         static_cast<const double *>(seeds), nullptr, 0, iterations, h, cutoff,
                         static_cast<double *>(out), static_cast<double *>(weight), 0, nullptr,
                         MsRun{});
    unsigned long long sums[kMsSums];
    ms_finish(s, sums);
    if (candidates)
      *candidates = sums[0];
    if (in_support)
      *in_support = sums[1];
    if (updates)
      *updates = sums[2];
  });
}

}  // extern "C"
