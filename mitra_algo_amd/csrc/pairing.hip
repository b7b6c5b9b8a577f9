// Curvature-signature sort and pruned sample-pair search.
//
// Restates src/cpu/pairing.cpp:14-35: every selected-ordinal pair a < b with
// |dkmin| <= t*M and |dkmax| <= t*M (M = max of the four |k| and 1e-10) and
// |p_a - p_b| >= 0.03*box, emitted in (a asc, b asc) order.
//   1. gather : selected ids -> ordinal-order sample records (AoS: one
//               64-byte line per FP32 sample, two per FP64 sample, so a
//               random access touches one or two lines) and a 32-bit
//               order-preserving key of kmin (-0 canonicalised);
//   2. sort   : LDS-staged stable LSD radix sort of (key, ordinal), 4 passes;
//   3. index  : the kmin-sorted samples are cut into buckets of 64, each
//               sorted by kmax inside one wavefront (64-lane bitonic network):
//               a two-level (kmin bucket, kmax) index;
//   4. windows: per source conservative intervals for kmin and kmax;
//   5. walk   : one wavefront per source; each lane binary-searches the kmax
//               interval in one bucket of the kmin interval (both bounds in
//               one loop), the wave concatenates the sub-ranges and tests
//               their candidates with coalesced loads.  Accepted ordinals
//               (b > a) are counted; up to 128 of them are put in CPU order
//               (64-lane bitonic network, or ranks in LDS beyond 64) and
//               parked in the source's slab;
//   6. scan   : exclusive prefix of the counts -> output offsets; the total is
//               read back once and checked against the budget before any
//               output is written;
//   7. place  : each slab is copied to its offset (coalesced); sources with
//               more than 128 pairs walk their targets in ordinal order in
//               4096-ordinal chunks (count, scan, fill), so order never
//               depends on a capacity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "pairing.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;  // sources per block
constexpr int kSlab = 128;                // pairs parked per source by the walk
constexpr uint64_t kLongChunk = 4096;     // ordinals per wave in the long path
constexpr int kBucket = 64;               // samples per kmin bucket
constexpr int kPivot = 8;                 // bucket entries per search segment
constexpr double kEps = 2.220446049250313e-16;
constexpr uint32_t kNone = 0xFFFFFFFFu;
// Signature sort key: the top kSigBits of the order-preserving 32-bit key of
// kmin (sign, exponent and 12 mantissa bits: classes 2^-12 of the value
// wide), sorted in three 8-bit passes instead of four.  Samples of one class
// keep their ordinal order.  Windows are searched on the same classes, so a
// window holds every sample of the classes its kmin interval touches: a
// superset of the exact window (every candidate is decided exactly), and the
// pairs and their order do not depend on the sorted order.
constexpr int kSigBits = 24;

__device__ inline uint32_t order_key32(double v) {
  if (v == 0.0)
    v = 0.0;  // -0 and +0 compare equal: one key
  uint64_t u = __double_as_longlong(v);
  u = (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
  return static_cast<uint32_t>(u >> 32);
}

__device__ inline float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1)
    v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}

// Flags: 1 = selected id out of range, 2 = non-finite coordinate/curvature.
// Batches: the centre is the sample's mesh centre and the sort key carries
// the mesh id above the 32-bit curvature key, so each mesh sorts apart.
__global__ void k_gather_samples(const double *x, const double *y, const double *z,
                                 const double *feat, uint64_t n, const int32_t *sel, uint64_t S,
                                 double cx, double cy, double cz, Segs sg, double *r64,
                                 float *r32, uint64_t *keys, uint32_t *vals, unsigned *flags,
                                 unsigned *rmax_bits) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = i < S;
  float r = 0.0f;
  if (live) {
    uint64_t mesh = 0;
    if (sg.seg) {
      mesh = sg.seg[i];
      cx = sg.center[3 * mesh], cy = sg.center[3 * mesh + 1], cz = sg.center[3 * mesh + 2];
    }
    int32_t id = sel[i];
    double v[kRec];
    for (int k = 0; k < kRec; ++k)
      v[k] = 0;
    if (id < 0 || static_cast<uint64_t>(id) >= n) {
      atomicOr(flags, 1u);
    } else {
      v[L_X] = x[id];
      v[L_Y] = y[id];
      v[L_Z] = z[id];
      v[L_K0] = feat[9 * n + id];
      v[L_K1] = feat[10 * n + id];
      for (int k = 0; k < 9; ++k)
        v[L_NX + k] = feat[k * n + id];
      bool fin = isfinite(v[L_X]) && isfinite(v[L_Y]) && isfinite(v[L_Z]) &&
                 isfinite(v[L_K0]) && isfinite(v[L_K1]);
      if (!fin) {
        atomicOr(flags, 2u);
        for (int k = 0; k < kRec; ++k)
          v[k] = 0;  // keep every later comparison well defined
      }
    }
    double2 *d2 = reinterpret_cast<double2 *>(r64 + i * kRec);
    for (int k = 0; k < kRec / 2; ++k)
      d2[k] = make_double2(v[2 * k], v[2 * k + 1]);
    if (r32) {
      float f[kRec];
      f[L_X] = static_cast<float>(v[L_X] - cx);
      f[L_Y] = static_cast<float>(v[L_Y] - cy);
      f[L_Z] = static_cast<float>(v[L_Z] - cz);
      for (int k = L_K0; k < kRec; ++k)
        f[k] = static_cast<float>(v[k]);
      float4 *f4 = reinterpret_cast<float4 *>(r32 + i * kRec);
      for (int k = 0; k < kRec / 4; ++k)
        f4[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
      r = fmaxf(fabsf(f[L_X]), fmaxf(fabsf(f[L_Y]), fabsf(f[L_Z])));
    }
    keys[i] = (mesh << kSigBits) | (order_key32(v[L_K0]) >> (32 - kSigBits));
    vals[i] = static_cast<uint32_t>(i);
  }
  // largest centre-relative coordinate (sets the FP32 guard bands); every
  // lane takes part in the reduction
  if (r32) {
    r = wave_max(r);
    if ((threadIdx.x & (kWave - 1)) == 0)
      atomicMax(rmax_bits, __float_as_uint(r));  // non-negative floats order as uints
  }
}

// One wavefront per bucket of 64 kmin-sorted samples: sort by (kmax, ordinal)
// with a 64-lane bitonic network and write the bucketed index, with each
// entry's curvatures and position beside it (the walk reads a candidate's
// fields with coalesced loads at its bucketed position, not at its ordinal).
__device__ inline void bucket_index_block(
    unsigned blk, const uint32_t *ord, const double *r64, const float *r32, uint64_t S,
                   uint32_t *bord, double2 *bk, float2 *bkf, double4 *pb, float4 *pbf,
                   double *piv, float *pivf) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t b = static_cast<uint64_t>(blk) * kWaves + threadIdx.x / kWave;
  const uint64_t nb = (S + kBucket - 1) / kBucket;
  if (b >= nb)
    return;
  const uint64_t j = b * kBucket + lane;
  uint32_t o = kNone;
  double key = INFINITY;
  if (j < S) {
    o = ord[j];
    key = r64[static_cast<uint64_t>(o) * kRec + L_K1];
  }
  for (int k = 2; k <= kWave; k <<= 1)
    for (int m = k >> 1; m > 0; m >>= 1) {
      double pk = __shfl_xor(key, m, kWave);
      uint32_t po = __shfl_xor(o, m, kWave);
      bool up = (lane & k) == 0;
      bool lower = (lane & m) == 0;
      bool mine_less = key < pk || (key == pk && o < po);
      bool keep = (lower == up) ? mine_less : !mine_less;
      if (!keep) {
        key = pk;
        o = po;
      }
    }
  // every kPivot-th kmax of the bucket (infinite past the samples), for a
  // search that reads them in one go before one segment of kPivot entries
  if (lane % kPivot == 0) {
    piv[b * (kBucket / kPivot) + lane / kPivot] = key;
    if (pivf)  // rounded as the FP32 record rounds it
      pivf[b * (kBucket / kPivot) + lane / kPivot] = static_cast<float>(key);
  }
  if (j < S) {
    const double *r = r64 + static_cast<uint64_t>(o) * kRec;
    bord[j] = o;
    bk[j] = make_double2(r[L_K0], key);
    pb[j] = make_double4(r[L_X], r[L_X + 1], r[L_X + 2], 0.0);
    if (bkf) {
      const float *f = r32 + static_cast<uint64_t>(o) * kRec;
      bkf[j] = make_float2(f[L_K0], f[L_K1]);
      pbf[j] = make_float4(f[L_X], f[L_X + 1], f[L_X + 2], 0.f);
    }
  }
}

// Conservative intervals.  If a pair passes, with Ma = max(|ka|,|Ka|,1e-10)
// and M the pair magnitude: M <= Ma/(1-t) (if the other sample dominates,
// its magnitude m satisfies m <= Ma + t m), hence |ka - kb| <= t Ma/(1-t)
// and |Ka - Kb| <= t Ma/(1-t).  t is inflated by 16 eps (the reference
// rounds both the difference and t*M), the radius by 16 eps and the
// endpoints by 16 eps of their magnitude; t >= .5 or t below the normal
// range use the sample's whole mesh.
// win: [0,S) first sorted position, [S,2S) end position; kw: kmax interval.
__device__ inline void windows_block(unsigned blk, const double *r64, const uint64_t *skeys,
                                     uint64_t S, double t, const uint32_t *seg, uint64_t *win,
                                     double2 *kw) {
  uint64_t i = static_cast<uint64_t>(blk) * kThreads + threadIdx.x;
  if (i >= S)
    return;
  const uint64_t mesh = seg ? static_cast<uint64_t>(seg[i]) << kSigBits : 0;
  uint32_t qlo = 0, qhi = 0xFFFFFFFFu;
  double klo = -INFINITY, khi = INFINITY;
  double tp = t * (1 + 16 * kEps);
  if (tp < 0.5 && t > 1e-290) {
    double a = r64[i * kRec + L_K0], b = r64[i * kRec + L_K1];
    double ma = fmax(fmax(fabs(a), fabs(b)), 1e-10);
    double r = tp * ma / (1 - tp) * (1 + 16 * kEps);
    double pa = 16 * kEps * (fabs(a) + r), pb = 16 * kEps * (fabs(b) + r);
    qlo = order_key32(a - r - pa);
    qhi = order_key32(a + r + pa);
    klo = b - r - pb;
    khi = b + r + pb;
  }
  const uint64_t kl = mesh | (qlo >> (32 - kSigBits)), kh = mesh | (qhi >> (32 - kSigBits));
  uint64_t f = 0, c = S;
  while (c) {  // first position with key >= kl
    uint64_t h = c / 2, m = f + h;
    if (skeys[m] < kl) {
      f = m + 1;
      c -= h + 1;
    } else
      c = h;
  }
  const uint64_t lo = f;
  f = 0;
  c = S;
  while (c) {  // first position with key > kh
    uint64_t h = c / 2, m = f + h;
    if (skeys[m] <= kh) {
      f = m + 1;
      c -= h + 1;
    } else
      c = h;
  }
  win[i] = lo;
  win[S + i] = f;
  kw[i] = make_double2(klo, khi);
}

// The bucket index (blocks [0, index_blocks)) and the windows (the rest)
// in one launch: both read the sorted order, neither the other's output.
__global__ void __launch_bounds__(kThreads)
    k_index_windows(unsigned index_blocks, const uint32_t *ord, const double *r64,
                    const float *r32, uint64_t S, uint32_t *bord, double2 *bk, float2 *bkf,
                    double4 *pb, float4 *pbf, double *piv, float *pivf, const uint64_t *skeys,
                    double t, const uint32_t *seg, uint64_t *win, double2 *kw) {
  if (blockIdx.x < index_blocks)
    bucket_index_block(blockIdx.x, ord, r64, r32, S, bord, bk, bkf, pb, pbf, piv, pivf);
  else
    windows_block(blockIdx.x - index_blocks, r64, skeys, S, t, seg, win, kw);
}

// Exact reference predicate in FP64 (pairing.cpp:22-28): the curvature
// test, then the distance test.
__device__ inline bool curv_ok64(double a0, double a1, double b0, double b1, double t) {
  double m = fabs(a0);
  double c1 = fabs(a1), c2 = fabs(b0), c3 = fabs(b1);
  m = m < c1 ? c1 : m;
  m = m < c2 ? c2 : m;
  m = m < c3 ? c3 : m;
  m = m < 1e-10 ? 1e-10 : m;
  double tol = t * m;
  return !(fabs(a0 - b0) > tol || fabs(a1 - b1) > tol);
}
__device__ inline bool dist_ok64(double ax, double ay, double az, double bx, double by, double bz,
                                 double dist) {
  double dx = ax - bx, dy = ay - by, dz = az - bz;
  return !(sqrt(dx * dx + dy * dy + dz * dz) < dist);
}
__device__ inline bool accept64(double a0, double a1, double b0, double b1, double ax, double ay,
                                double az, double bx, double by, double bz, double t,
                                double dist) {
  double m = fabs(a0);
  double c1 = fabs(a1), c2 = fabs(b0), c3 = fabs(b1);
  m = m < c1 ? c1 : m;
  m = m < c2 ? c2 : m;
  m = m < c3 ? c3 : m;
  m = m < 1e-10 ? 1e-10 : m;
  double tol = t * m;
  if (fabs(a0 - b0) > tol || fabs(a1 - b1) > tol)
    return false;
  double dx = ax - bx, dy = ay - by, dz = az - bz;
  return !(sqrt(dx * dx + dy * dy + dz * dz) < dist);
}

// FP32 screen with guard bands: 0 reject, 1 accept, 2 undecided (FP64 recheck).
__device__ inline int screen32(float a0, float a1, float b0, float b1, float t, float gk) {
  float m = fmaxf(fmaxf(fabsf(a0), fabsf(a1)), fmaxf(fmaxf(fabsf(b0), fabsf(b1)), 1e-10f));
  if (!(m < 1e30f))
    return 2;
  float tol = t * m;
  float d0 = fabsf(a0 - b0), d1 = fabsf(a1 - b1);
  float hi = tol * (1.0f + gk), lo = tol * (1.0f - gk);
  if (d0 > hi || d1 > hi)
    return 0;
  if (d0 <= lo && d1 <= lo)
    return 1;
  return 2;
}
__device__ inline int screen32_dist(float dx, float dy, float dz, float dist, float gd) {
  float d = sqrtf(dx * dx + dy * dy + dz * dz);
  if (d < dist * (1.0f - gd))
    return 0;
  if (d >= dist * (1.0f + gd))
    return 1;
  return 2;
}

struct PairArgs {
  const double *r64;
  const float *r32;
  const uint32_t *bord;
  const double2 *bk;   // (kmin, kmax) at bucketed positions, kmax sorted per bucket
  const float2 *bkf;   // FP32 copy
  const double4 *pb;   // positions (x, y, z, 0) at bucketed positions
  const float4 *pbf;   // FP32 copy
  const double *piv;   // per bucket: the kmax of every kPivot-th entry
  const float *pivf;   // FP32 copy
  const uint64_t *win;
  const double2 *kw;
  uint64_t S;
  double t, dist;
  float t32, dist32, gk, gd;
  // distance band = gd0 + gd1 * (largest FP32 coordinate), set per launch
  float gd0, gd1;
  const unsigned *rmax_bits;
  uint32_t *cand, *cand2, *rechk;  // per source statistics
  Segs sg;                         // batches: per-mesh box and ordinal range
};

template <bool kGuarded> __device__ inline PairArgs with_bands(const PairArgs &in) {
  PairArgs A = in;
  if (kGuarded)
    A.gd = in.gd0 + in.gd1 * __uint_as_float(*in.rmax_bits);
  return A;
}

struct Source {
  uint64_t i;     // the source's ordinal
  double a0, a1, ax, ay, az;  // FP64 fields (the FP64 walk only)
  float f0, f1, fx, fy, fz;
  double dist;    // 0.03 x the mesh's box size
  float dist32;
  uint64_t ohi;   // end of the source's mesh in ordinals
};

// Minimum pair distance and ordinal end of source i's mesh.
__device__ inline void source_mesh(const PairArgs &A, uint64_t i, double &dist, uint64_t &ohi) {
  if (A.sg.seg) {
    const uint32_t m = A.sg.seg[i];
    dist = 0.03 * A.sg.box[m];
    ohi = A.sg.ostart[m + 1];
  } else {
    dist = A.dist;
    ohi = A.S;
  }
}

// Curvatures of bucketed position j: FP32 for the guarded walk, FP64 for
// the plain one (loaded beside the ordinal, not after it).
template <bool kGuarded> struct BucketK;
template <> struct BucketK<true> {
  using T = float2;
  using P = float4;
  __device__ static T load(const PairArgs &A, uint64_t j) { return A.bkf[j]; }
  __device__ static P pos(const PairArgs &A, uint64_t j) { return A.pbf[j]; }
};
template <> struct BucketK<false> {
  using T = double2;
  using P = double4;
  __device__ static T load(const PairArgs &A, uint64_t j) { return A.bk[j]; }
  __device__ static P pos(const PairArgs &A, uint64_t j) { return A.pb[j]; }
};

// Decide candidate at bucketed position j (ordinal o, curvatures bkj,
// position pj: loaded with the ordinal, not after the curvature test).
template <bool kGuarded>
__device__ inline bool decide(const PairArgs &A, const Source &src, uint64_t j, uint32_t o,
                              const typename BucketK<kGuarded>::T &bkj,
                              const typename BucketK<kGuarded>::P &pj, unsigned &rechecks) {
  if constexpr (!kGuarded) {
    const double2 b = bkj;
    double m = fmax(fmax(fabs(src.a0), fabs(src.a1)), fmax(fmax(fabs(b.x), fabs(b.y)), 1e-10));
    double tol = A.t * m;
    if (fabs(src.a0 - b.x) > tol || fabs(src.a1 - b.y) > tol)
      return false;
    const double4 p = pj;
    return accept64(src.a0, src.a1, b.x, b.y, src.ax, src.ay, src.az, p.x, p.y, p.z, A.t,
                    src.dist);
  } else {
    const float2 b = bkj;
    int r = screen32(src.f0, src.f1, b.x, b.y, A.t32, A.gk);
    if (r == 0)
      return false;
    if (r == 1) {
      const float4 p = pj;
      int d = screen32_dist(src.fx - p.x, src.fy - p.y, src.fz - p.z, src.dist32, A.gd);
      if (d != 2)
        return d == 1;
    }
    // FP64 recheck: the source's FP64 fields are read here, not kept live
    // through the walk, curvatures first and positions after (rechecks are
    // rare; registers are not)
    ++rechecks;
    const double *rs = A.r64 + src.i * kRec;
    const double2 bb = A.bk[j];
    if (!curv_ok64(rs[L_K0], rs[L_K1], bb.x, bb.y, A.t))
      return false;
    const double4 q = A.pb[j];
    return dist_ok64(rs[L_X], rs[L_Y], rs[L_Z], q.x, q.y, q.z, src.dist);
  }
}

// Two sources per wave, one per 32-lane half: a source's window spans ~17
// buckets and ~19 tested candidates (config 2), so a whole wave per source
// leaves most lanes idle in both loops.
constexpr int kHalf = kWave / 2;

// Per-source LDS for concatenating the bucket sub-ranges.
struct HalfRanges {
  uint32_t start[kHalf];
  uint32_t pre[kHalf + 1];
};

// Lanes of my half below me in a wave-wide ballot, and the half's count.
__device__ inline uint32_t half_below(uint64_t mask, int lane) {
  const uint32_t below = static_cast<uint32_t>(lanes_below(mask));
  return lane < kHalf ? below : below - static_cast<uint32_t>(__popc(static_cast<uint32_t>(mask)));
}
__device__ inline uint32_t half_count(uint64_t mask, int lane) {
  return static_cast<uint32_t>(__popc(static_cast<uint32_t>(lane < kHalf ? mask : mask >> 32)));
}
// The larger of the two halves' values (uniform over the wave).
__device__ inline uint64_t halves_max(uint64_t v) {
  const uint64_t w = static_cast<uint64_t>(__shfl_xor(static_cast<long long>(v), kHalf, kWave));
  return v > w ? v : w;
}

// A source's window and kmax interval: loaded one source ahead in the
// walk, so they arrive beside the current source's first loads (the
// bucket pivots depend on them; the source's record does not).
struct SrcPre {
  uint64_t jlo, jhi;
  double2 kw;
};
__device__ inline SrcPre preload_source(const PairArgs &A, uint64_t ii) {
  SrcPre r;
  r.jlo = A.win[ii];
  r.jhi = A.win[A.S + ii];
  r.kw = A.kw[ii];
  return r;
}

// Lower bound fl = entries with kmax < lo and upper bound fu = entries with
// kmax <= hi of the kmax-sorted bucket at [base, base + nb), by counting:
// the kPivot pivots (every kPivot-th entry's kmax) locate each bound's
// segment, whose other entries are then counted; two rounds of independent
// loads instead of one dependent load per bisection step.  K = float: the
// FP32 copies of kmax (rounded to nearest, so still sorted) against bounds
// rounded outwards, which counts at most as many below lo and at least as
// many up to hi: the interval holds the exact one.
__device__ inline float f_down(double x) {
  const float f = static_cast<float>(x);
  return static_cast<double>(f) > x ? nextafterf(f, -INFINITY) : f;
}
__device__ inline float f_up(double x) {
  const float f = static_cast<float>(x);
  return static_cast<double>(f) < x ? nextafterf(f, INFINITY) : f;
}
template <class K, class K2>
__device__ inline void bucket_bounds(const K *pv_in, const K2 *kk, uint64_t base, uint64_t nb,
                                     K lo, K hi, uint32_t &fl, uint32_t &fu) {
  constexpr int kSeg = kBucket / kPivot;
  K pv[kSeg];
  for (int q = 0; q < kSeg; ++q)
    pv[q] = pv_in[q];
  int sl = 0, su = 0;  // pivots below the lower bound / not above the upper
  for (int q = 0; q < kSeg; ++q) {
    sl += pv[q] < lo;
    su += !(pv[q] > hi);
  }
  // segment s - 1 holds the bound (entries 1 .. kPivot - 1 of it are
  // counted; its pivot already was); s = 0: the bound is 0
  K el[kPivot - 1], eu[kPivot - 1];
  const uint64_t bl = base + static_cast<uint64_t>(sl > 0 ? sl - 1 : 0) * kPivot;
  const uint64_t bu = base + static_cast<uint64_t>(su > 0 ? su - 1 : 0) * kPivot;
  for (int q = 0; q < kPivot - 1; ++q) {
    el[q] = bl + 1 + q < base + nb ? kk[bl + 1 + q].y : K(INFINITY);
    eu[q] = bu + 1 + q < base + nb ? kk[bu + 1 + q].y : K(INFINITY);
  }
  fl = 0, fu = 0;
  if (sl > 0) {
    fl = static_cast<uint32_t>((sl - 1) * kPivot + 1);
    for (int q = 0; q < kPivot - 1; ++q)
      fl += bl + 1 + q < base + nb && el[q] < lo;
  }
  if (su > 0) {
    fu = static_cast<uint32_t>((su - 1) * kPivot + 1);
    for (int q = 0; q < kPivot - 1; ++q)
      fu += bu + 1 + q < base + nb && !(eu[q] > hi);
  }
}

// Walks every candidate of source i in this lane's half (live: the half has
// a source); f(accepted, ordinal) is called with all lanes active once per
// candidate chunk of both halves.
template <bool kGuarded, class F>
__device__ inline void walk_source(const PairArgs &A, uint64_t i, bool live, const SrcPre &pre,
                                   HalfRanges &W, int lane, unsigned &rechecks, uint64_t &cand2,
                                   F &&f) {
  const uint64_t S = A.S;
  const int hl = lane & (kHalf - 1);
  const uint64_t ii = live ? i : 0;  // loads stay in range for an idle half (pre: for ii)
  const uint64_t jlo = pre.jlo, jhi = pre.jhi;
  const bool has = live && jlo < jhi;
  const double2 kw = pre.kw;
  Source src;
  src.i = ii;
  {
    if (!kGuarded) {
      const double2 *p = reinterpret_cast<const double2 *>(A.r64 + ii * kRec);
      double2 p0 = p[0], p1 = p[1], p2 = p[2];
      src.ax = p0.x, src.ay = p0.y, src.az = p1.x, src.a0 = p1.y, src.a1 = p2.x;
    } else {
      const float4 *q = reinterpret_cast<const float4 *>(A.r32 + ii * kRec);
      float4 q0 = q[0], q1 = q[1];
      src.fx = q0.x, src.fy = q0.y, src.fz = q0.z, src.f0 = q0.w, src.f1 = q1.x;
    }
    source_mesh(A, ii, src.dist, src.ohi);
    src.dist32 = static_cast<float>(src.dist);
  }
  const uint64_t b0 = jlo / kBucket;
  const uint64_t nbk = has ? (jhi - 1) / kBucket - b0 + 1 : 0;  // buckets of the window
  const uint64_t nbk_max = halves_max(nbk);
  for (uint64_t off = 0; off < nbk_max; off += kHalf) {
    const uint64_t b = b0 + off + static_cast<uint64_t>(hl);
    uint32_t u = 0, len = 0;
    if (off + static_cast<uint64_t>(hl) < nbk) {
      const uint64_t base = b * kBucket;
      const uint64_t nb = S - base < static_cast<uint64_t>(kBucket) ? S - base : kBucket;
      uint32_t fl = 0, fu = 0;
      if constexpr (kGuarded)  // FP32 keys and widened bounds: a superset
        bucket_bounds(A.pivf + b * (kBucket / kPivot), A.bkf, base, nb, f_down(kw.x),
                      f_up(kw.y), fl, fu);
      else
        bucket_bounds(A.piv + b * (kBucket / kPivot), A.bk, base, nb, kw.x, kw.y, fl, fu);
      // an infinite bound also counts the infinite pivots past a short
      // bucket's end
      fl = fl < nb ? fl : static_cast<uint32_t>(nb);
      fu = fu < nb ? fu : static_cast<uint32_t>(nb);
      u = static_cast<uint32_t>(base + fl);
      len = fu > fl ? static_cast<uint32_t>(fu - fl) : 0u;
    }
    // exclusive prefix of the sub-range lengths within each half: one scan
    // over the wave, the second half less the first half's total
    const uint32_t xw = wave_incl_scan(len);
    const uint32_t t0 = __builtin_amdgcn_readlane(xw, kHalf - 1);
    const uint32_t t1 = __builtin_amdgcn_readlane(xw, kWave - 1) - t0;
    const uint32_t x = lane < kHalf ? xw : xw - t0;
    const uint32_t total = lane < kHalf ? t0 : t1;
    W.start[hl] = u;
    W.pre[hl] = x - len;
    if (hl == 0)
      W.pre[kHalf] = total;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    cand2 += total;
    const uint32_t tmax = t0 > t1 ? t0 : t1;
    for (uint32_t c0 = 0; c0 < tmax; c0 += kHalf) {
      const uint32_t c = c0 + static_cast<uint32_t>(hl);
      bool ok = false;
      uint32_t o = 0;
      if (c < total) {
        // owner sub-range: last l with pre[l] <= c
        int l = 0;
        for (int step = kHalf / 2; step > 0; step >>= 1)
          if (W.pre[l + step] <= c)
            l += step;
        const uint64_t j = W.start[l] + (c - W.pre[l]);
        o = A.bord[j];
        const typename BucketK<kGuarded>::T bkj = BucketK<kGuarded>::load(A, j);
        const typename BucketK<kGuarded>::P pj = BucketK<kGuarded>::pos(A, j);
        if (o != kNone && o > i && o < src.ohi)
          ok = decide<kGuarded>(A, src, j, o, bkj, pj, rechecks);
      }
      f(ok, o);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Single walk per source: counts every accepted pair and parks up to kSlab
// of them, in ordinal order, in the source's slab.  Two sources per wave
// (one per 32-lane half).  quota == 0: wave w of the grid walks sources 2w
// and 2w + 1.  quota > 0 (persistent): the resident waves claim `quota`
// sources at a time from the queue counter until every source is taken, so
// sources of uneven cost (and meshes of uneven size) balance across the
// chip; every wave's last claim returns a start past the end, so all exit.
template <bool kGuarded>
__global__ void __launch_bounds__(kThreads, kGuarded ? 5 : 1)  // guarded: <= 96 registers
    k_pair_walk(PairArgs Ain, uint64_t *counts, uint32_t *slab, uint32_t *long_list,
                unsigned *long_count, unsigned quota, unsigned *queue) {
  __shared__ HalfRanges ranges[2 * kWaves];
  __shared__ uint32_t list[2 * kWaves][kSlab];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int half = lane / kHalf, hl = lane & (kHalf - 1);
  const PairArgs A = with_bands<kGuarded>(Ain);
  HalfRanges &W = ranges[2 * wave + half];
  uint32_t *L = list[2 * wave + half];
  // walks source i in this lane's half (live: the half has one)
  auto walk_one = [&](uint64_t i, bool live, const SrcPre &pre) {
    uint32_t cnt = 0;
    uint64_t cand2 = 0;
    unsigned rechecks = 0;
    walk_source<kGuarded>(A, i, live, pre, W, lane, rechecks, cand2, [&](bool ok, uint32_t o) {
      const uint64_t mask = ballot(ok);
      const uint32_t pos = cnt + half_below(mask, lane);
      if (ok && pos < static_cast<uint32_t>(kSlab))
        L[pos] = o;
      cnt += half_count(mask, lane);
    });
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // partners in ordinal order: each lane places its entries by rank (the
    // number of smaller ordinals; ordinals are distinct)
    if (live && cnt <= static_cast<uint32_t>(kSlab))
      for (uint32_t e = hl; e < cnt; e += kHalf) {
        const uint32_t v = L[e];
        uint32_t r = 0;
        for (uint32_t k = 0; k < cnt; ++k)
          r += L[k] < v;
        slab[i * kSlab + r] = v;
      }
    for (int o = kHalf / 2; o > 0; o >>= 1)  // the half's recheck count
      rechecks += static_cast<unsigned>(__shfl_xor(static_cast<int>(rechecks), o, kWave));
    if (live && hl == 0) {
      counts[i] = cnt;
      if (cnt > static_cast<uint32_t>(kSlab)) {
        unsigned s = atomicAdd(long_count, 1u);
        long_list[s] = static_cast<uint32_t>(i);
      }
      uint64_t w = pre.jhi - pre.jlo;
      A.cand[i] = static_cast<uint32_t>(w < kNone ? w : kNone);
      A.cand2[i] = static_cast<uint32_t>(cand2 < kNone ? cand2 : kNone);
      A.rechk[i] = rechecks;
    }
    // the next sources rewrite this wave's lists and ranges
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  };
  // source of pair slot q in this lane's half (sample 0 stands in past S)
  auto source_of = [&](uint64_t q) -> uint64_t {
    const uint64_t i = 2 * q + static_cast<uint64_t>(half);
    return i < A.S ? i : 0;
  };
  if (quota == 0 && !kGuarded) {  // one pair slot per wave
    const uint64_t q = static_cast<uint64_t>(blockIdx.x) * kWaves + wave;
    const uint64_t i = 2 * q + static_cast<uint64_t>(half);
    if (!ballot(i < A.S))
      return;
    walk_one(i, i < A.S, preload_source(A, source_of(q)));
    return;
  }
  if (quota == 0) {
    // static (FP32-guarded walk): wave w of the grid walks pair slots w,
    // w + waves, ... (two sources each), the next slot's window loading
    // while the current one is walked
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * kWaves;
    const uint64_t slots = (A.S + 1) / 2;
    uint64_t q = static_cast<uint64_t>(blockIdx.x) * kWaves + wave;
    if (q >= slots)
      return;  // the whole wave
    SrcPre cur = preload_source(A, source_of(q));
    for (; q < slots; q += waves) {
      const uint64_t i = 2 * q + static_cast<uint64_t>(half);
      SrcPre nxt = cur;
      if (q + waves < slots)
        nxt = preload_source(A, source_of(q + waves));
      walk_one(i, i < A.S, cur);
      cur = nxt;
    }
    return;
  }
  for (;;) {
    unsigned base = 0;
    if (lane == 0)
      base = atomicAdd(queue, quota);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, kWave));
    if (base >= A.S)
      break;
    const uint64_t end = base + quota < A.S ? base + quota : A.S;
    for (uint64_t t = base; t < end; t += 2) {
      const uint64_t i = t + static_cast<uint64_t>(half);
      walk_one(i, i < end, preload_source(A, i < end ? i : 0));
    }
  }
}

// Copies every slab (count <= kSlab) to its output offset, kPlaceLanes
// lanes per source (a source parks ~8 pairs at config 2).
// Writes stay below `cap` (the pair capacity), also when a graph replay
// meets counts that differ from the ones it was captured with.
constexpr int kPlaceLanes = 8;
__global__ void __launch_bounds__(kThreads)
    k_pair_place(uint64_t S, const uint64_t *counts, const uint64_t *offsets,
                 const uint32_t *slab, const int32_t *sel, int32_t *out_a, int32_t *out_b,
                 uint64_t cap) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const uint64_t i = t / kPlaceLanes;
  const uint32_t sub = static_cast<uint32_t>(t % kPlaceLanes);
  if (i >= S)
    return;
  const uint64_t c = counts[i];
  if (c > static_cast<uint64_t>(kSlab))
    return;  // written by the long path
  const uint64_t o0 = offsets[i];
  const int32_t src = sel ? sel[i] : static_cast<int32_t>(i);
  // every parked pair, also beyond the first kPlaceLanes (slabs hold up to kSlab)
  for (uint64_t e = sub; e < c; e += kPlaceLanes) {
    const uint32_t o = slab[i * kSlab + e];
    const uint64_t p = o0 + e;
    if (p >= cap)
      break;
    out_a[p] = src;
    out_b[p] = sel ? sel[o] : static_cast<int32_t>(o);
  }
}

// Sources with more than kSlab pairs walk their targets in ordinal order,
// split into chunks of kLongChunk ordinals (one wave per (source, chunk)):
// a counting pass, a scan over the flattened (source, chunk) counts, and a
// writing pass that places each chunk's pairs after the earlier chunks'.
__device__ inline void long_range(uint64_t i, uint64_t c, uint64_t ohi, uint64_t &lo,
                                  uint64_t &hi) {
  lo = c * kLongChunk;
  hi = lo + kLongChunk < ohi ? lo + kLongChunk : ohi;
  if (lo < i + 1)
    lo = i + 1;
}

__global__ void __launch_bounds__(kThreads)
    k_long_count(PairArgs A, const uint32_t *long_list, uint64_t long_count, uint64_t chunks,
                 uint32_t *chunk_counts) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / kWave;
  if (w >= long_count * chunks)
    return;
  const uint64_t i = long_list[w / chunks], c = w % chunks;
  double dist;
  uint64_t ohi, lo, hi;
  source_mesh(A, i, dist, ohi);
  long_range(i, c, ohi, lo, hi);
  const double *ri = A.r64 + i * kRec;
  const double a0 = ri[L_K0], a1 = ri[L_K1], ax = ri[L_X], ay = ri[L_Y], az = ri[L_Z];
  uint32_t cnt = 0;
  for (uint64_t b0 = lo; b0 < hi; b0 += kWave) {
    uint64_t b = b0 + lane;
    bool ok = false;
    if (b < hi) {
      const double *rb = A.r64 + b * kRec;
      ok = accept64(a0, a1, rb[L_K0], rb[L_K1], ax, ay, az, rb[L_X], rb[L_Y], rb[L_Z], A.t,
                    dist);
    }
    cnt += static_cast<uint32_t>(__popcll(ballot(ok)));
  }
  if (lane == 0)
    chunk_counts[w] = cnt;
}

__global__ void __launch_bounds__(kThreads)
    k_long_fill(PairArgs A, const uint32_t *long_list, uint64_t long_count, uint64_t chunks,
                const uint64_t *chunk_off, const uint64_t *offsets, const int32_t *sel,
                int32_t *out_a, int32_t *out_b, uint64_t cap) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / kWave;
  if (w >= long_count * chunks)
    return;
  const uint64_t src_w = w / chunks, c = w % chunks;
  const uint64_t i = long_list[src_w];
  double dist;
  uint64_t ohi, lo, hi;
  source_mesh(A, i, dist, ohi);
  long_range(i, c, ohi, lo, hi);
  const double *ri = A.r64 + i * kRec;
  const double a0 = ri[L_K0], a1 = ri[L_K1], ax = ri[L_X], ay = ri[L_Y], az = ri[L_Z];
  const int32_t src = sel ? sel[i] : static_cast<int32_t>(i);
  uint64_t cursor = offsets[i] + (chunk_off[w] - chunk_off[src_w * chunks]);
  const uint64_t end = offsets[i + 1] < cap ? offsets[i + 1] : cap;
  for (uint64_t b0 = lo; b0 < hi; b0 += kWave) {
    uint64_t b = b0 + lane;
    bool ok = false;
    if (b < hi) {
      const double *rb = A.r64 + b * kRec;
      ok = accept64(a0, a1, rb[L_K0], rb[L_K1], ax, ay, az, rb[L_X], rb[L_Y], rb[L_Z], A.t,
                    dist);
    }
    uint64_t mask = ballot(ok);
    uint64_t p = cursor + lanes_below(mask);
    if (ok && p < end) {
      out_a[p] = src;
      out_b[p] = sel ? sel[b] : static_cast<int32_t>(b);
    }
    cursor += __popcll(mask);
  }
}

// Sets *mismatch when the device bytes differ from the recorded ones.
__global__ void k_check_equal(const unsigned char *a, const unsigned char *b, uint64_t bytes,
                              unsigned *mismatch) {
  bool diff = false;
  for (uint64_t i = threadIdx.x; i < bytes; i += blockDim.x)
    diff = diff || a[i] != b[i];
  if (diff)
    atomicOr(mismatch, 1u);
}

__global__ void k_sum_stats(const uint32_t *a, const uint32_t *b, const uint32_t *c, uint64_t n,
                            unsigned long long *out) {
  __shared__ unsigned long long s[3][kThreads];
  unsigned long long x = 0, y = 0, z = 0;
  for (uint64_t i = threadIdx.x; i < n; i += kThreads) {
    x += a[i];
    y += b[i];
    z += c[i];
  }
  s[0][threadIdx.x] = x;
  s[1][threadIdx.x] = y;
  s[2][threadIdx.x] = z;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; ++k)
        s[k][threadIdx.x] += s[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3)
    out[threadIdx.x] = s[threadIdx.x][0];
}

// -------------------------------------------------------------- prune
__global__ void k_prune_flags(const double *kmin, const double *kmax, const uint8_t *valid,
                              uint64_t n, const int32_t *sample, uint64_t ns, double threshold,
                              uint32_t *flag, unsigned *bad) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ns)
    return;
  int32_t v = sample[i];
  if (v < 0 || static_cast<uint64_t>(v) >= n) {
    atomicOr(bad, 1u);
    flag[i] = 0;
    return;
  }
  double a = fabs(kmin[v]), b = fabs(kmax[v]);
  double m = a < b ? b : a;
  flag[i] = valid[v] && m > 1e-10 && fabs(kmin[v] - kmax[v]) / m > threshold;
}
__global__ void k_compact(const int32_t *sample, const uint32_t *flag, const uint64_t *off,
                          uint64_t ns, int32_t *out) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < ns && flag[i])
    out[off[i]] = sample[i];
}

}  // namespace

void read_back(Session *s, void *dst, const void *src, size_t bytes) {
  if (s->replay == Replay::kCapture) {
    // a captured step reads nothing back: it takes the recorded value and
    // leaves a device check that the replayed value is the same
    if (s->rec_pos >= s->rec.size() || s->rec[s->rec_pos].size() != bytes)
      raise(MVS_EINVAL, "graph capture diverged from the recorded step");
    std::memcpy(dst, s->rec[s->rec_pos].data(), bytes);
    MVS_LAUNCH(s, "graph_check", k_check_equal, 1, kThreads, 0,
               static_cast<const unsigned char *>(src), s->rec_dev.ptr + s->rec_off[s->rec_pos],
               static_cast<uint64_t>(bytes), s->mismatch_dev);
    ++s->rec_pos;
    return;
  }
  void *h = host_stage(s, bytes);
  MVS_HIP(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, s->stream));
  note_copy(false, bytes);
  MVS_HIP(hipStreamSynchronize(s->stream));
  std::memcpy(dst, h, bytes);
  if (s->replay == Replay::kRecord) {
    const unsigned char *b = static_cast<const unsigned char *>(dst);
    s->rec.emplace_back(b, b + bytes);
  }
}

// Guard-band widths of the FP32 screen (relative).  Curvature: inputs rounded
// once to FP32 and three FP32 operations give |err(d)| <= 4 eps32 M and
// |err(tol)| <= 3 eps32 t M, i.e. (4/t + 3) eps32 relative to tol; the band
// adds a factor 4 and a 1e-4 floor.  Distance: with every FP32 coordinate
// bounded by rmax (measured during the gather), the computed distance
// carries at most ~8 eps32 rmax of absolute error; the band is
// 1e-4 + 16 eps32 rmax / dist (set on the device, see with_bands).
static float guard_curvature(double t) {
  return static_cast<float>(1e-4 + 16.0 * 5.960464477539063e-08 * (1.0 + 1.0 / t));
}

struct PairState {
  double *r64;
  float *r32;
  uint32_t *ord, *bord, *slab;
  double2 *bk, *kw;
  float2 *bkf;
  double4 *pb;
  float4 *pbf;
  double *piv;
  float *pivf;
  uint64_t *keys, *win, *counts, *offsets;
  uint32_t *long_list, *cand, *cand2, *rechk;
};

static PairState pair_state(Session *s, uint64_t S, bool guarded) {
  PairState st{};
  const uint64_t Sp = S ? S : 1;
  st.r64 = scratch_as<double>(s, kSlotSamples, kRec * Sp);
  st.r32 = guarded ? scratch_as<float>(s, kSlotPlanSamples32, kRec * Sp) : nullptr;
  st.ord = scratch_as<uint32_t>(s, kSlotPlanOrd, 2 * Sp);  // [0,S) sort input, [S,2S) sorted
  st.bk = scratch_as<double2>(s, kSlotPlanSortedK, Sp);
  st.bkf = guarded ? scratch_as<float2>(s, kSlotPlanSortedK32, Sp) : nullptr;
  st.pb = scratch_as<double4>(s, kSlotPlanPosB, Sp);
  st.pbf = guarded ? scratch_as<float4>(s, kSlotPlanPosB32, Sp) : nullptr;
  {  // FP64 pivots, then (guarded) their FP32 copy
    const uint64_t np = (Sp + kBucket - 1) / kBucket * (kBucket / kPivot);
    st.piv = scratch_as<double>(s, kSlotPlanPivots, np + (np + 1) / 2);
    st.pivf = guarded ? reinterpret_cast<float *>(st.piv + np) : nullptr;
  }
  st.bord = scratch_as<uint32_t>(s, kSlotWindows, Sp);
  st.slab = scratch_as<uint32_t>(s, kSlotPlanSlab, kSlab * Sp);
  st.keys = scratch_as<uint64_t>(s, kSlotPlanKeys, 2 * Sp);
  st.win = scratch_as<uint64_t>(s, kSlotPlanWin, 2 * Sp);
  st.kw = scratch_as<double2>(s, kSlotPlanKw, Sp);
  st.counts = scratch_as<uint64_t>(s, kSlotPlanCounts, Sp + 1);
  st.offsets = scratch_as<uint64_t>(s, kSlotPlanOffsets, Sp + 1);
  st.long_list = scratch_as<uint32_t>(s, kSlotPlanLong, Sp + 1);
  uint32_t *stats = scratch_as<uint32_t>(s, kSlotPlanStats, 3 * Sp);
  st.cand = stats;
  st.cand2 = stats + Sp;
  st.rechk = stats + 2 * Sp;
  return st;
}

// box: the mesh's box size, or the smallest box of a batch (guard bands).
static PairArgs pair_args(const PairState &st, uint64_t S, double t, double box, bool guarded,
                          const unsigned *flags, const Segs &segs) {
  PairArgs A{};
  A.r64 = st.r64;
  A.r32 = st.r32;
  A.bord = st.bord;
  A.bk = st.bk;
  A.bkf = st.bkf;
  A.pb = st.pb;
  A.pbf = st.pbf;
  A.piv = st.piv;
  A.pivf = st.pivf;
  A.win = st.win;
  A.kw = st.kw;
  A.S = S;
  A.t = t;
  A.dist = 0.03 * box;
  A.t32 = static_cast<float>(t);
  A.dist32 = static_cast<float>(A.dist);
  A.gk = guarded ? guard_curvature(t) : 0.f;
  A.gd = 0.f;
  A.gd0 = 1e-4f;
  A.gd1 = static_cast<float>(16.0 * 5.960464477539063e-08 / A.dist);
  A.rmax_bits = flags + kRmaxWord;
  A.cand = st.cand;
  A.cand2 = st.cand2;
  A.rechk = st.rechk;
  A.sg = segs;
  return A;
}

const uint64_t *plan_pair_offsets(Session *s) {
  return scratch_as<uint64_t>(s, kSlotPlanOffsets, (s->pair_plan.S ? s->pair_plan.S : 1) + 1);
}

const double *plan_samples64(Session *s) {
  return scratch_as<double>(s, kSlotSamples, kRec * (s->pair_plan.S ? s->pair_plan.S : 1));
}
const float *plan_samples32(Session *s) {
  return scratch_as<float>(s, kSlotPlanSamples32, kRec * (s->pair_plan.S ? s->pair_plan.S : 1));
}

void pair_count_device(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                       const double *feat, uint64_t S, const int32_t *sel, double box, double t,
                       uint64_t max_pairs, int precision, const double *center, uint64_t *P_out,
                       const Segs &segs) {
  if (precision != MVS_FP64 && precision != MVS_FP32_GUARDED)
    raise(MVS_EINVAL, "unknown precision");
  if (!(box > 0) || !std::isfinite(box) || !(t > 0 && t <= 1))
    raise(MVS_EINVAL, "pairing requires positive finite box size and threshold in (0,1]");
  if (S > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
    raise(MVS_EOVERFLOW, "selected ordinals exceed int32 range");
  PairPlan &plan = s->pair_plan;
  plan = PairPlan{};
  plan.S = S;
  plan.threshold = t;
  plan.box = box;
  plan.precision = precision;
  plan.selected = sel;
  plan.segs = segs;
  const bool guarded = precision == MVS_FP32_GUARDED;
  PairState st = pair_state(s, S, guarded);
  unsigned *flags = plan_flags(s);
  // the pairing words and the vote phase's out-of-range flag (kVoteBadWord)
  static_assert(kVoteBadWord < 22, "the flag words cleared here");
  zero_device(s, flags, 22 * sizeof(unsigned));
  double cx = center ? center[0] : 0, cy = center ? center[1] : 0, cz = center ? center[2] : 0;
  if (S)
    MVS_LAUNCH(s, "pair_gather", k_gather_samples, blocks_for(S, kThreads), kThreads, 0, x, y, z,
               feat, n, sel, S, cx, cy, cz, segs, st.r64, st.r32, st.keys, st.ord, flags,
               flags + kRmaxWord);
  if (S < 2) {  // no pair possible; like the reference, nothing to validate
    plan.P = 0;
    plan.valid = true;
    *P_out = 0;
    return;
  }
  int key_bits = kSigBits;  // batches: the mesh id sits above the curvature key
  if (segs.seg)
    while (key_bits < 64 && (s->batch_meshes >> (key_bits - kSigBits)) != 0)
      ++key_bits;
  radix_sort_pairs_u64(s, st.keys, st.ord, st.keys + S, st.ord + S, S, key_bits);
  const uint64_t nb = (S + kBucket - 1) / kBucket;
  const unsigned index_blocks = blocks_for(nb, kWaves);
  MVS_LAUNCH(s, "pair_index", k_index_windows, index_blocks + blocks_for(S, kThreads), kThreads,
             0, index_blocks, st.ord + S, st.r64, st.r32, S, st.bord, st.bk, st.bkf, st.pb,
             st.pbf, st.piv, st.pivf, st.keys + S, t, segs.seg, st.win, st.kw);
  unsigned *long_count = flags + 1;
  unsigned *queue = flags + 12;  // cleared with the flags above
  PairArgs A = pair_args(st, S, t, box, guarded, flags, segs);
  const unsigned quota = s->tune.walk_quota;
  // two sources per wave; the FP32-guarded walk and the persistent one
  // (quota > 0) loop on a fixed set of blocks (static slots: the blocks the
  // device holds at once; claims from the queue: the tuned count), the FP64
  // walk takes one pair slot per wave
  unsigned grid = blocks_for(S, 2 * kWaves);
  if (quota) {
    grid = std::min<unsigned>(grid, s->tune.walk_blocks);
  } else if (guarded) {
    if (!s->walk_resident) {
      int per_cu = 0, cus = 0;
      MVS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pair_walk<true>, kThreads,
                                                           0));
      MVS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
      s->walk_resident = static_cast<unsigned>(std::max(1, per_cu) * std::max(1, cus));
    }
    grid = std::min<unsigned>(grid, s->walk_resident);
  }
  if (guarded)
    MVS_LAUNCH(s, "pair_walk", k_pair_walk<true>, grid, kThreads, 0, A, st.counts, st.slab,
               st.long_list, long_count, quota, queue);
  else
    MVS_LAUNCH(s, "pair_walk", k_pair_walk<false>, grid, kThreads, 0, A, st.counts, st.slab,
               st.long_list, long_count, quota, queue);
  // flags, long-source count and total read back in one synchronisation
  // (the scan writes the total beside the flags too)
  uint64_t *tot = reinterpret_cast<uint64_t *>(flags + 2);
  exclusive_scan_u64(s, st.counts, st.offsets, S, tot);
  unsigned host[4];
  read_back(s, host, flags, sizeof host);
  if (host[0] & 1u)
    raise(MVS_EINVAL, "selected vertex index out of range");
  if (host[0] & 2u)
    raise(MVS_EINVAL, "pairing requires finite coordinates and curvatures");
  uint64_t P;
  std::memcpy(&P, host + 2, sizeof P);
  if (P > max_pairs || P > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
    raise(MVS_EBUDGET, "Pair budget exceeded; lower --sample-ratio or tighten --pair-threshold");
  plan.P = P;
  plan.long_sources = host[1];
  plan.valid = true;
  *P_out = P;
}

void pair_fill_device(Session *s, int32_t *a, int32_t *b, bool ordinals) {
  PairPlan &plan = s->pair_plan;
  if (!plan.valid)
    raise(MVS_EINVAL, "pair fill without a preceding count");
  if (plan.P == 0)
    return;
  const uint64_t S = plan.S;
  const bool guarded = plan.precision == MVS_FP32_GUARDED;
  PairState st = pair_state(s, S, guarded);
  PairArgs A = pair_args(st, S, plan.threshold, plan.box, guarded, plan_flags(s), plan.segs);
  const int32_t *sel = ordinals ? nullptr : plan.selected;
  MVS_LAUNCH(s, "pair_place", k_pair_place, blocks_for(S, kThreads / kPlaceLanes), kThreads, 0,
             S, st.counts,
             st.offsets, st.slab, sel, a, b, plan.P);
  if (plan.long_sources) {
    const uint64_t L = plan.long_sources, chunks = (S + kLongChunk - 1) / kLongChunk;
    uint32_t *cc = scratch_as<uint32_t>(s, kSlotPlanLongCnt, L * chunks);
    uint64_t *co = scratch_as<uint64_t>(s, kSlotPlanLongOff, L * chunks + 1);
    unsigned grid = blocks_for(L * chunks, kWaves);
    MVS_LAUNCH(s, "pair_long", k_long_count, grid, kThreads, 0, A, st.long_list, L, chunks, cc);
    exclusive_scan_u32(s, cc, co, L * chunks);
    MVS_LAUNCH(s, "pair_long", k_long_fill, grid, kThreads, 0, A, st.long_list, L, chunks, co,
               st.offsets, sel, a, b, plan.P);
  }
}

void pair_statistics(Session *s, uint64_t out[3]) {
  PairPlan &plan = s->pair_plan;
  out[0] = out[1] = out[2] = 0;
  if (!plan.valid || plan.S < 2)
    return;
  PairState st = pair_state(s, plan.S, plan.precision == MVS_FP32_GUARDED);
  unsigned long long *dst = scratch_as<unsigned long long>(s, kSlotBlock, 4);
  MVS_LAUNCH(s, "pair_stats", k_sum_stats, 1, kThreads, 0, st.cand, st.cand2, st.rechk, plan.S,
             dst);
  unsigned long long h[3];
  read_back(s, h, dst, sizeof h);
  out[0] = h[0], out[1] = h[1], out[2] = h[2];
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_pair_count(mvs_session *sp, uint64_t n, const double *x, const double *y,
                       const double *z, const double *feat, uint64_t S, const int32_t *sel,
                       double box, const double center[3], double t, uint64_t max_pairs,
                       int precision, uint64_t *pairs) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!pairs || !center)
      raise(MVS_EINVAL, "null argument");
    pair_count_device(s, n, x, y, z, feat, S, sel, box, t, max_pairs, precision, center, pairs);
  });
}

int mvs_dev_pair_fill(mvs_session *sp, int32_t *a, int32_t *b, int ordinals) {
  return guarded([&] { pair_fill_device(as_session(sp), a, b, ordinals != 0); });
}

int mvs_dev_pair_stats(mvs_session *sp, uint64_t *window_candidates, uint64_t *index_candidates,
                       uint64_t *fp64_rechecks) {
  return guarded([&] {
    uint64_t st[3];
    pair_statistics(as_session(sp), st);
    if (window_candidates)
      *window_candidates = st[0];
    if (index_candidates)
      *index_candidates = st[1];
    if (fp64_rechecks)
      *fp64_rechecks = st[2];
  });
}

int mvs_prune(mvs_session *sp, const mvs_features *f, uint64_t ns, const int32_t *sample,
              double prune, int32_t *out, uint64_t *count) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!f || !count || (ns && (!sample || !out)))
      raise(MVS_EINVAL, "null argument");
    *count = 0;
    if (ns == 0)
      return;
    uint64_t n = f->count;
    DBuf<double> k(2 * (n ? n : 1));
    DBuf<uint8_t> valid(n ? n : 1);
    DBuf<int32_t> dsample(ns), dout(ns);
    DBuf<uint32_t> flag(ns);
    DBuf<uint64_t> off(ns + 1);
    DBuf<unsigned> bad(1);
    copy_h2d(s, k.ptr, f->lane[9], n * 8);
    copy_h2d(s, k.ptr + n, f->lane[10], n * 8);
    copy_h2d(s, valid.ptr, f->valid, n);
    copy_h2d(s, dsample.ptr, sample, ns * 4);
    zero_device(s, bad.ptr, sizeof(unsigned));
    MVS_LAUNCH(s, "prune_flags", k_prune_flags, blocks_for(ns, 256), 256, 0, k.ptr, k.ptr + n,
               valid.ptr, n, dsample.ptr, ns, 1 - prune, flag.ptr, bad.ptr);
    exclusive_scan_u32(s, flag.ptr, off.ptr, ns);
    MVS_LAUNCH(s, "prune_compact", k_compact, blocks_for(ns, 256), 256, 0, dsample.ptr, flag.ptr,
               off.ptr, ns, dout.ptr);
    unsigned hbad = 0;
    copy_d2h(s, &hbad, bad.ptr, sizeof hbad);
    if (hbad)
      raise(MVS_EINVAL, "sample vertex index out of range");
    uint64_t kept = 0;
    copy_d2h(s, &kept, off.ptr + ns, sizeof kept);
    copy_d2h(s, out, dout.ptr, kept * 4);
    *count = kept;
  });
}

int mvs_dev_prune(mvs_session *sp, uint64_t n, const double *feat, const uint8_t *valid,
                  uint64_t ns, const int32_t *sample, double prune, int32_t *out,
                  uint64_t *count) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!count || (ns && (!feat || !valid || !sample || !out)))
      raise(MVS_EINVAL, "null argument");
    *count = 0;
    if (ns == 0)
      return;
    uint32_t *flag = scratch_as<uint32_t>(s, kSlotFlags, ns);
    uint64_t *off = scratch_as<uint64_t>(s, kSlotOffsets, ns + 2);
    unsigned *bad = reinterpret_cast<unsigned *>(off + ns + 1);  // past the scan's total
    zero_device(s, bad, sizeof(unsigned));
    MVS_LAUNCH(s, "prune_flags", k_prune_flags, blocks_for(ns, 256), 256, 0, feat + 9 * n,
               feat + 10 * n, valid, n, sample, ns, 1 - prune, flag, bad);
    exclusive_scan_u32(s, flag, off, ns);
    MVS_LAUNCH(s, "prune_compact", k_compact, blocks_for(ns, 256), 256, 0, sample, flag, off, ns,
               out);
    struct {
      unsigned bad;
      uint64_t kept;
    } h{};
    copy_d2h(s, &h.bad, bad, sizeof h.bad);
    copy_d2h(s, &h.kept, off + ns, sizeof h.kept);
    if (h.bad)
      raise(MVS_EINVAL, "sample vertex index out of range");
    *count = h.kept;
  });
}

int mvs_pairing(mvs_session *sp, const mvs_points *v, const mvs_features *f, uint64_t S,
                const int32_t *selected, double box, double t, uint64_t max_pairs, int precision,
                uint64_t *count) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!v || !f || !count || (S && !selected))
      raise(MVS_EINVAL, "null argument");
    *count = 0;
    s->out_count = 0;
    if (S < 2)
      return;
    uint64_t n = v->count;
    if (f->count != n)
      raise(MVS_EINVAL, "SoA component lengths differ");
    DBuf<double> dv(3 * (n ? n : 1)), feat(11 * (n ? n : 1));
    DBuf<int32_t> dsel(S);
    copy_h2d(s, dv.ptr, v->x, n * 8);
    copy_h2d(s, dv.ptr + n, v->y, n * 8);
    copy_h2d(s, dv.ptr + 2 * n, v->z, n * 8);
    for (int k = 0; k < 11; ++k)
      copy_h2d(s, feat.ptr + k * n, f->lane[k], n * 8);
    copy_h2d(s, dsel.ptr, selected, S * 4);
    // FP32 coordinates are taken relative to the bounding-box midpoint
    double ctr[3] = {0, 0, 0};
    if (n) {
      const double *c3[3] = {v->x, v->y, v->z};
      for (int k = 0; k < 3; ++k) {
        double lo = c3[k][0], hi = c3[k][0];
        for (uint64_t i = 1; i < n; ++i) {
          lo = c3[k][i] < lo ? c3[k][i] : lo;
          hi = hi < c3[k][i] ? c3[k][i] : hi;
        }
        ctr[k] = std::isfinite(lo) && std::isfinite(hi) ? 0.5 * (lo + hi) : 0.0;
      }
    }
    uint64_t P = 0;
    pair_count_device(s, n, dv.ptr, dv.ptr + n, dv.ptr + 2 * n, feat.ptr, S, dsel.ptr, box, t,
                      max_pairs, precision, ctr, &P);
    s->out_a.ensure(P ? P : 1);
    s->out_b.ensure(P ? P : 1);
    pair_fill_device(s, s->out_a.ptr, s->out_b.ptr, false);
    sync(s);
    s->out_count = P;
    *count = P;
  });
}

int mvs_pairing_fetch(mvs_session *sp, int32_t *a, int32_t *b) {
  return guarded([&] {
    Session *s = as_session(sp);
    copy_d2h(s, a, s->out_a.ptr, s->out_count * 4);
    copy_d2h(s, b, s->out_b.ptr, s->out_count * 4);
  });
}

}  // extern "C"
