// Per-vertex curvature tensor and principal frames (FP64).
//
// Restates src/cpu/curvature.cpp:14-79 for gfx950:
//   * mean edge length: a fixed-order tree sum that computes the lengths as
//     it sums them (deterministic; it agrees with the reference's sequential
//     sum to a few ulps);
//   * gather pass: one thread per vertex computes and sums its interior-edge
//     contributions d d^T beta len / mean (curvature.cpp:21-40).  The
//     reference adds them in edge order; edges are sorted by (lo, hi), so a
//     vertex's incident edges in edge order are exactly its ring neighbours
//     in ascending order (edges (u, v) with u < v first, then (v, w) with
//     w > v).  The gather walks the sorted ring, which gives the same
//     summation order without an incident-edge CSR, and recomputes each
//     edge's contribution at both endpoints (same operations, same bits)
//     rather than writing 9 doubles per edge and reading them back twice;
//   * smoothing: ping-pong (T_i + sum_nbr T_j) / (deg + 1) (curvature.cpp:41-51);
//   * frames: tangent-plane 2x2 eigen-decomposition (curvature.cpp:54-77).
// Tensors are 9-lane SoA so every lane access is coalesced.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;

// Edge length of edge i (curvature.cpp:17-18).
__device__ inline double edge_length(const double *x, const double *y, const double *z,
                                     const mvs_edge &E) {
  return norm(v3(x[E.hi], y[E.hi], z[E.hi]) - v3(x[E.lo], y[E.lo], z[E.lo]));
}

// Fixed-shape tree reduction: each block sums a tile of kSumItems values per
// thread in a fixed order (thread t adds items t, t + 256, ... of the tile),
// then a tree over the threads; a second launch sums the block partials the
// same way.  The first level computes the edge lengths it sums (no length
// array), issuing every load of a thread before the first sum.
constexpr int kSumItems = 8;
constexpr uint64_t kSumTile = kThreads * kSumItems;

template <bool kLengths>
__global__ void __launch_bounds__(kThreads)
    k_sum_tiles(const double *in, const double *x, const double *y, const double *z,
                const mvs_edge *edges, uint64_t n, double *out) {
  __shared__ double s[kThreads];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSumTile + threadIdx.x;
  double val[kSumItems];
  if constexpr (kLengths) {
    int32_t hi[kSumItems], lo[kSumItems];
    for (int r = 0; r < kSumItems; ++r) {
      const uint64_t i = base + r * kThreads;
      hi[r] = i < n ? edges[i].hi : 0;
      lo[r] = i < n ? edges[i].lo : 0;
    }
    for (int r = 0; r < kSumItems; ++r)
      val[r] = base + r * kThreads < n
                   ? norm(v3(x[hi[r]], y[hi[r]], z[hi[r]]) - v3(x[lo[r]], y[lo[r]], z[lo[r]]))
                   : 0.0;
  } else {
    for (int r = 0; r < kSumItems; ++r)
      val[r] = base + r * kThreads < n ? in[base + r * kThreads] : 0.0;
  }
  double acc = 0;
  for (int r = 0; r < kSumItems; ++r)
    if (base + r * kThreads < n)
      acc += val[r];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[blockIdx.x] = s[0];
}

__global__ void k_mean_finish(const double *sum, uint64_t ne, double *mean) {
  *mean = sum[0] / static_cast<double>(ne);
}

// estart[v] = index of the first edge with lo == v (count of edges with lo < v).
__global__ void k_greater_count(const uint64_t *off, const int32_t *idx, uint64_t n,
                                uint32_t *cnt) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  uint32_t c = 0;
  for (uint64_t j = off[v]; j < off[v + 1]; ++j)
    c += idx[j] > static_cast<int32_t>(v);
  cnt[v] = c;
}

// Per vertex, the dihedral contributions d d^T beta len / mean of its
// interior edges (curvature.cpp:21-37), summed in edge order, then divided
// by their count (curvature.cpp:38-40).  Edges are sorted by (lo, hi), so
// the vertex's edges in edge order are its ring neighbours in ascending
// order: for a neighbour u > v the edge is the next one of v's own block
// (estart[v], ...); for u < v it is the edge of u's block whose hi is v.
// Each edge's contribution is computed by both of its endpoints, with the
// same operations, instead of being stored and gathered.
//
// The ring is taken kGatherChunk slots at a time, phase by phase (ring entries,
// edge ids, edge records, then positions and face normals), so a thread has
// a chunk's loads in flight together instead of one dependent chain per
// slot; the sums still run in slot order.
constexpr int kChunk = 8;        // smoothing
constexpr int kGatherChunk = 4;  // gather (more registers per slot)

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_gather(const uint64_t *__restrict__ off, const int32_t *__restrict__ idx,
             const uint64_t *__restrict__ estart, const mvs_edge *__restrict__ edges, uint64_t ne,
             const double *__restrict__ x, const double *__restrict__ y,
             const double *__restrict__ z, const double *__restrict__ fn, uint64_t nf,
             const double *__restrict__ mean, uint64_t n, double *__restrict__ T) {
  uint64_t v = static_cast<uint64_t>(xcd_block(blocks_for(n, kThreads))) * kThreads + threadIdx.x;
  if (v >= n)
    return;
  const double m = *mean;
  double acc[9];
  for (int k = 0; k < 9; ++k)
    acc[k] = 0;
  int count = 0;
  const int32_t self = static_cast<int32_t>(v);
  uint64_t up = estart[v];
  const uint64_t up_end = estart[v + 1] < ne ? estart[v + 1] : ne;
  const uint64_t j1 = off[v + 1];
  for (uint64_t j0 = off[v]; j0 < j1; j0 += kGatherChunk) {
    const int cnt = j1 - j0 < kGatherChunk ? static_cast<int>(j1 - j0) : kGatherChunk;
    int32_t u[kGatherChunk];
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r)
      u[r] = r < cnt ? idx[j0 + r] : self;
    // edge ids: upper neighbours in order from v's block; lower ones from
    // their own block (its first entries read together)
    uint64_t e[kGatherChunk], b0[kGatherChunk], b1[kGatherChunk];
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r) {
      e[r] = ne;
      b0[r] = b1[r] = 0;
      if (r < cnt) {
        if (u[r] > self) {
          e[r] = up < up_end ? up : ne;
          ++up;
        } else {
          b0[r] = estart[u[r]];
          b1[r] = estart[u[r] + 1] < ne ? estart[u[r] + 1] : ne;
        }
      }
    }
    int32_t ehi[kGatherChunk];
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r)
      ehi[r] = e[r] < ne ? edges[e[r]].hi : (b0[r] < b1[r] ? edges[b0[r]].hi : -1);
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r) {
      if (r >= cnt)
        continue;
      if (u[r] > self) {
        if (e[r] < ne && ehi[r] != u[r])
          e[r] = ne;
      } else if (b0[r] < b1[r] && ehi[r] == self) {
        e[r] = b0[r];
      } else {
        for (uint64_t k = b0[r] + 1; k < b1[r]; ++k)
          if (edges[k].hi == self) {
            e[r] = k;
            break;
          }
      }
    }
    mvs_edge E[kGatherChunk];
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r)
      E[r] = e[r] < ne ? edges[e[r]] : mvs_edge{0, 0, 0, 0, 0, -1};
#pragma unroll
    for (int r = 0; r < kGatherChunk; ++r) {
      if (r >= cnt || E[r].face1 < 0)  // no edge, or a boundary edge
        continue;
      ++count;
      V3 d = v3(x[E[r].b], y[E[r].b], z[E[r].b]) - v3(x[E[r].a], y[E[r].a], z[E[r].a]);
      const double len = norm(d);
      d = unit(d);
      const int32_t f0 = E[r].face0, f1 = E[r].face1;
      const V3 n1 = v3(fn[f0], fn[nf + f0], fn[2 * nf + f0]);
      const V3 n2 = v3(fn[f1], fn[nf + f1], fn[2 * nf + f1]);
      double c = dot(n1, n2);
      c = c < -1.0 ? -1.0 : (1.0 < c ? 1.0 : c);
      const double beta = atan2(dot(cross(n1, n2), d), c);
      // d_i d_k and d_k d_i are the same product, so the tensor is
      // symmetric to the bit: six values serve the nine lanes
      const double dd[3] = {d.x, d.y, d.z};
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = i; k < 3; ++k) {
          const double t = dd[i] * dd[k] * beta * len / m;
          acc[i * 3 + k] += t;
          if (k != i)
            acc[k * 3 + i] += t;
        }
    }
  }
  const int c = count > 1 ? count : 1;
  for (int k = 0; k < 9; ++k)
    T[k * n + v] = acc[k] / c;
}

// One smoothing pass: (T_v + sum of the ring's T_u in ring order) / (deg + 1)
// (curvature.cpp:41-51).  Lane by lane, a chunk of the ring's values is
// loaded together and then added in order.
__global__ void __launch_bounds__(kThreads)
    k_smooth(const uint64_t *__restrict__ off, const int32_t *__restrict__ idx, uint64_t n,
             const double *__restrict__ T, double *__restrict__ out) {
  uint64_t v = static_cast<uint64_t>(xcd_block(blocks_for(n, kThreads))) * kThreads + threadIdx.x;
  if (v >= n)
    return;
  double acc[9];
  for (int k = 0; k < 9; ++k)
    acc[k] = T[k * n + v];
  const uint64_t j1 = off[v + 1], j0s = off[v];
  for (uint64_t j0 = j0s; j0 < j1; j0 += kChunk) {
    const int cnt = j1 - j0 < kChunk ? static_cast<int>(j1 - j0) : kChunk;
    uint64_t u[kChunk];
#pragma unroll
    for (int r = 0; r < kChunk; ++r)
      u[r] = r < cnt ? static_cast<uint64_t>(idx[j0 + r]) : v;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double t[kChunk];
#pragma unroll
      for (int r = 0; r < kChunk; ++r)
        t[r] = T[k * n + u[r]];
#pragma unroll
      for (int r = 0; r < kChunk; ++r)
        if (r < cnt)
          acc[k] += t[r];
    }
  }
  double deg1 = static_cast<double>(j1 - j0s + 1);
  for (int k = 0; k < 9; ++k)
    out[k * n + v] = acc[k] / deg1;
}

__global__ void k_frames(const double *T, const double *normals, const uint64_t *off, uint64_t n,
                         double *feat, uint8_t *valid) {
  uint64_t i = static_cast<uint64_t>(xcd_block(blocks_for(n, kThreads))) * kThreads + threadIdx.x;
  if (i >= n)
    return;
  double vals[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint8_t ok = 0;
  V3 nn = v3(normals[i], normals[n + i], normals[2 * n + i]);
  if (!(norm(nn) < 0.5 || off[i + 1] - off[i] < 3)) {
    V3 axis = fabs(nn.x) < 0.8 ? v3(1, 0, 0) : v3(0, 1, 0);
    V3 u = unit(cross(nn, axis)), w = cross(nn, u);
    M3 t;
    for (int k = 0; k < 9; ++k)
      t.m[k] = T[k * n + i];
    double a = dot(u, apply(t, u)), b = dot(u, apply(t, w)), c = dot(w, apply(t, w));
    double angle = 0.5 * atan2(2 * b, a - c);
    V3 high = u * cos(angle) + w * sin(angle);
    V3 low = unit(cross(high, nn));
    double delta = hypot(a - c, 2 * b);
    double kmin = (a + c - delta) / 2, kmax = (a + c + delta) / 2;
    V3 dmax = unit(cross(nn, low));
    double r[11] = {nn.x, nn.y, nn.z, low.x, low.y, low.z, dmax.x, dmax.y, dmax.z, kmin, kmax};
    for (int k = 0; k < 11; ++k)
      vals[k] = r[k];
    ok = isfinite(kmin) && isfinite(kmax) && delta > 1e-10;
  }
  for (int k = 0; k < 11; ++k)
    feat[k * n + i] = vals[k];
  valid[i] = ok;
}

}  // namespace

void curvature_device(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                      const double *normals, uint64_t nf, const double *fn, const uint64_t *off,
                      const int32_t *idx, uint64_t ne, const mvs_edge *edges, int smoothing,
                      double *feat, uint8_t *valid) {
  if (n == 0)
    return;
  if (smoothing < 0)
    raise(MVS_EINVAL, "smoothing must be non-negative");
  double *Ta = scratch_as<double>(s, kSlotMisc2, 9 * n);
  double *Tb = scratch_as<double>(s, kSlotMisc3, 9 * n);
  uint32_t *gcnt = scratch_as<uint32_t>(s, kSlotCounts, n);
  uint64_t *estart = scratch_as<uint64_t>(s, kSlotOffsets, n + 1);
  // partial sums: level one over the edges, then levels over the partials
  // until one value is left
  const uint64_t tiles = blocks_for(ne ? ne : 1, kSumTile);
  uint64_t total = 0;
  for (uint64_t c = tiles;; c = blocks_for(c, kSumTile)) {
    total += c;
    if (c == 1)
      break;
  }
  double *partial = scratch_as<double>(s, kSlotBlock, total + 1);
  double *mean = partial + total;
  if (ne) {
    MVS_LAUNCH(s, "curv_sum", k_sum_tiles<true>, static_cast<unsigned>(tiles), kThreads, 0,
               nullptr, x, y, z, edges, ne, partial);
    double *cur = partial;
    for (uint64_t c = tiles; c > 1;) {
      const uint64_t nb = blocks_for(c, kSumTile);
      MVS_LAUNCH(s, "curv_sum", k_sum_tiles<false>, static_cast<unsigned>(nb), kThreads, 0, cur,
                 nullptr, nullptr, nullptr, nullptr, c, cur + c);
      cur += c;
      c = nb;
    }
    MVS_LAUNCH(s, "curv_mean", k_mean_finish, 1, 1, 0, cur, ne, mean);
  }
  MVS_LAUNCH(s, "curv_greater", k_greater_count, blocks_for(n, kThreads), kThreads, 0, off, idx, n,
             gcnt);
  exclusive_scan_u32(s, gcnt, estart, n);
  const unsigned vgrid = xcd_grid(blocks_for(n, kThreads));
  MVS_LAUNCH(s, "curv_gather", k_gather, vgrid, kThreads, 0, off, idx, estart,
             edges, ne, x, y, z, fn, nf, mean, n, Ta);
  double *cur = Ta, *nxt = Tb;
  for (int it = 0; it < smoothing; ++it) {
    MVS_LAUNCH(s, "curv_smooth", k_smooth, vgrid, kThreads, 0, off, idx, n, cur,
               nxt);
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  MVS_LAUNCH(s, "curv_frames", k_frames, vgrid, kThreads, 0, cur, normals, off,
             n, feat, valid);
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_curvature(mvs_session *sp, uint64_t n, const double *x, const double *y,
                      const double *z, const double *normals3, uint64_t faces,
                      const double *face_normals3, const uint64_t *nbr_off,
                      const int32_t *nbr_idx, uint64_t edges, const mvs_edge *edge, int smoothing,
                      double *feat, uint8_t *valid) {
  return guarded([&] {
    Session *s = as_session(sp);
    curvature_device(s, n, x, y, z, normals3, faces, face_normals3, nbr_off, nbr_idx, edges, edge,
                     smoothing, feat, valid);
  });
}

int mvs_curvature(mvs_session *sp, const mvs_points *v, const mvs_topology *t, int smoothing,
                  const mvs_features_out *out) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!v || !t || !out)
      raise(MVS_EINVAL, "null argument");
    uint64_t n = v->count;
    if (t->normals.count != n)
      raise(MVS_EINVAL, "SoA component lengths differ");
    if (n == 0)
      return;
    uint64_t nf = t->face_normals.count, ne = t->edge_count;
    uint64_t nnz = t->neighbor_offsets[n];
    // the kernels index with these ids: reject inconsistent topology up front
    if (t->neighbor_offsets[0] != 0)
      raise(MVS_EINVAL, "Invalid CSR offsets");
    for (uint64_t i = 0; i < n; ++i)
      if (t->neighbor_offsets[i + 1] < t->neighbor_offsets[i])
        raise(MVS_EINVAL, "Invalid CSR offsets");
    for (uint64_t j = 0; j < nnz; ++j)
      if (t->neighbor_indices[j] < 0 || static_cast<uint64_t>(t->neighbor_indices[j]) >= n)
        raise(MVS_EINVAL, "neighbour index out of range");
    for (uint64_t e = 0; e < ne; ++e) {
      const mvs_edge &E = t->edges[e];
      auto in_v = [&](int32_t v) { return v >= 0 && static_cast<uint64_t>(v) < n; };
      auto in_f = [&](int32_t f) { return f >= 0 && static_cast<uint64_t>(f) < nf; };
      if (!in_v(E.lo) || !in_v(E.hi) || !in_v(E.a) || !in_v(E.b) || !in_f(E.face0) ||
          (E.face1 >= 0 && !in_f(E.face1)))
        raise(MVS_EINVAL, "edge record out of range");
    }
    DBuf<double> dv(3 * n), dn(3 * n), dfn(3 * (nf ? nf : 1)), feat(11 * n);
    DBuf<uint64_t> doff(n + 1);
    DBuf<int32_t> didx(nnz ? nnz : 1);
    DBuf<mvs_edge> de(ne ? ne : 1);
    DBuf<uint8_t> dvalid(n);
    copy_h2d(s, dv.ptr, v->x, n * 8);
    copy_h2d(s, dv.ptr + n, v->y, n * 8);
    copy_h2d(s, dv.ptr + 2 * n, v->z, n * 8);
    copy_h2d(s, dn.ptr, t->normals.x, n * 8);
    copy_h2d(s, dn.ptr + n, t->normals.y, n * 8);
    copy_h2d(s, dn.ptr + 2 * n, t->normals.z, n * 8);
    if (nf) {
      copy_h2d(s, dfn.ptr, t->face_normals.x, nf * 8);
      copy_h2d(s, dfn.ptr + nf, t->face_normals.y, nf * 8);
      copy_h2d(s, dfn.ptr + 2 * nf, t->face_normals.z, nf * 8);
    }
    copy_h2d(s, doff.ptr, t->neighbor_offsets, (n + 1) * 8);
    copy_h2d(s, didx.ptr, t->neighbor_indices, nnz * 4);
    copy_h2d(s, de.ptr, t->edges, ne * sizeof(mvs_edge));
    curvature_device(s, n, dv.ptr, dv.ptr + n, dv.ptr + 2 * n, dn.ptr, nf, dfn.ptr, doff.ptr,
                     didx.ptr, ne, de.ptr, smoothing, feat.ptr, dvalid.ptr);
    for (int k = 0; k < 11; ++k)
      copy_d2h(s, out->lane[k], feat.ptr + k * n, n * 8);
    copy_d2h(s, out->valid, dvalid.ptr, n);
    sync(s);
  });
}

}  // extern "C"
