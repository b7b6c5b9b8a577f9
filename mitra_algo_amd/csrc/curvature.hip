// Per-vertex curvature tensor and principal frames (FP64).
//
// Restates src/cpu/curvature.cpp:14-79 for gfx950:
//   * edge pass: one thread per edge computes the dihedral contribution
//     d d^T beta len / mean once (curvature.cpp:21-37);
//   * gather pass: one thread per vertex sums its interior-edge
//     contributions.  The reference adds them in edge order; edges are sorted
//     by (lo, hi), so a vertex's incident edges in edge order are exactly its
//     ring neighbours in ascending order (edges (u, v) with u < v first, then
//     (v, w) with w > v).  The gather walks the sorted ring, which gives the
//     same summation order without an incident-edge CSR;
//   * smoothing: ping-pong (T_i + sum_nbr T_j) / (deg + 1) (curvature.cpp:41-51);
//   * frames: tangent-plane 2x2 eigen-decomposition (curvature.cpp:54-77).
// Tensors are 9-lane SoA so every lane access is coalesced.
// The mean edge length is a fixed-order tree sum (deterministic); it agrees
// with the reference's sequential sum to a few ulps.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;

__global__ void k_edge_length(const double *x, const double *y, const double *z,
                              const mvs_edge *edges, uint64_t ne, double *len) {
  uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne)
    return;
  int32_t h = edges[e].hi, l = edges[e].lo;
  len[e] = norm(v3(x[h], y[h], z[h]) - v3(x[l], y[l], z[l]));
}

// Fixed-shape tree reduction: each block sums a contiguous tile in a fixed
// order; a second launch sums the block partials the same way.
__global__ void k_sum_tiles(const double *in, uint64_t n, uint64_t tile, double *out) {
  __shared__ double s[kThreads];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * tile;
  double acc = 0;
  for (uint64_t i = base + threadIdx.x; i < base + tile && i < n; i += kThreads)
    acc += in[i];
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

// Contribution of each interior edge (9 lanes, stride ne).  Boundary edges
// get a NaN marker in lane 0 so the gather can skip them without re-reading
// the edge record.
__global__ void k_edge_tensor(const double *x, const double *y, const double *z,
                              const double *fn, uint64_t nf, const mvs_edge *edges, uint64_t ne,
                              const double *mean, double *contrib) {
  uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= ne)
    return;
  mvs_edge E = edges[e];
  if (E.face1 < 0) {
    contrib[e] = __builtin_nan("");
    return;
  }
  V3 d = v3(x[E.b], y[E.b], z[E.b]) - v3(x[E.a], y[E.a], z[E.a]);
  double len = norm(d);
  d = unit(d);
  V3 n1 = v3(fn[E.face0], fn[nf + E.face0], fn[2 * nf + E.face0]);
  V3 n2 = v3(fn[E.face1], fn[nf + E.face1], fn[2 * nf + E.face1]);
  double c = dot(n1, n2);
  c = c < -1.0 ? -1.0 : (1.0 < c ? 1.0 : c);
  double beta = atan2(dot(cross(n1, n2), d), c);
  double m = *mean;
  double dd[3] = {d.x, d.y, d.z};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      contrib[(i * 3 + j) * ne + e] = dd[i] * dd[j] * beta * len / m;
}

// Edge id of (lo, hi): edges are sorted by (lo, hi), so it is the first edge
// with this lo plus the rank of hi among lo's ring neighbours greater than lo.
// Returns ne (no edge) when hi is not a greater neighbour of lo, which a
// consistent topology never produces.
__device__ inline uint64_t edge_id(const uint64_t *off, const int32_t *idx,
                                   const uint64_t *edge_start, int32_t lo, int32_t hi,
                                   uint64_t ne) {
  uint64_t r0 = off[lo], r1 = off[lo + 1];
  uint64_t first_greater = r0, pos = r1;
  for (uint64_t j = r0; j < r1; ++j) {
    int32_t u = idx[j];
    if (u <= lo)
      first_greater = j + 1;
    if (u == hi)
      pos = j;
  }
  if (pos == r1 || pos < first_greater)
    return ne;
  uint64_t e = edge_start[lo] + (pos - first_greater);
  return e < ne ? e : ne;
}

// edge_start[v] = index of the first edge with lo == v (count of edges with lo < v).
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

__global__ void k_gather(const uint64_t *off, const int32_t *idx, const uint64_t *edge_start,
                         const double *contrib, uint64_t ne, uint64_t n, double *T) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  double acc[9];
  for (int k = 0; k < 9; ++k)
    acc[k] = 0;
  int count = 0;
  int32_t self = static_cast<int32_t>(v);
  for (uint64_t j = off[v]; j < off[v + 1]; ++j) {
    int32_t u = idx[j];
    uint64_t e = u < self ? edge_id(off, idx, edge_start, u, self, ne)
                          : edge_id(off, idx, edge_start, self, u, ne);
    if (e >= ne)
      continue;
    double c0 = contrib[e];
    if (c0 != c0)  // boundary edge
      continue;
    ++count;
    acc[0] += c0;
    for (int k = 1; k < 9; ++k)
      acc[k] += contrib[k * ne + e];
  }
  int c = count > 1 ? count : 1;
  for (int k = 0; k < 9; ++k)
    T[k * n + v] = acc[k] / c;
}

__global__ void k_smooth(const uint64_t *off, const int32_t *idx, uint64_t n, const double *T,
                         double *out) {
  uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n)
    return;
  double acc[9];
  for (int k = 0; k < 9; ++k)
    acc[k] = T[k * n + v];
  for (uint64_t j = off[v]; j < off[v + 1]; ++j) {
    uint64_t u = static_cast<uint64_t>(idx[j]);
    for (int k = 0; k < 9; ++k)
      acc[k] += T[k * n + u];
  }
  double deg1 = static_cast<double>(off[v + 1] - off[v] + 1);
  for (int k = 0; k < 9; ++k)
    out[k * n + v] = acc[k] / deg1;
}

__global__ void k_frames(const double *T, const double *normals, const uint64_t *off, uint64_t n,
                         double *feat, uint8_t *valid) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
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
  double *len = scratch_as<double>(s, kSlotMisc0, ne ? ne : 1);
  double *contrib = scratch_as<double>(s, kSlotMisc1, 9 * (ne ? ne : 1));
  double *Ta = scratch_as<double>(s, kSlotMisc2, 9 * n);
  double *Tb = scratch_as<double>(s, kSlotMisc3, 9 * n);
  uint32_t *gcnt = scratch_as<uint32_t>(s, kSlotCounts, n);
  uint64_t *estart = scratch_as<uint64_t>(s, kSlotOffsets, n + 1);
  const uint64_t tile = 8192;
  unsigned tiles = blocks_for(ne ? ne : 1, tile);
  double *partial = scratch_as<double>(s, kSlotBlock, tiles + 1);
  double *mean = partial + tiles;
  if (ne) {
    MVS_LAUNCH(s, "curv_edge_length", k_edge_length, blocks_for(ne, kThreads), kThreads, 0, x, y,
               z, edges, ne, len);
    MVS_LAUNCH(s, "curv_sum", k_sum_tiles, tiles, kThreads, 0, len, ne, tile, partial);
    if (tiles > 1)
      MVS_LAUNCH(s, "curv_sum", k_sum_tiles, 1, kThreads, 0, partial, static_cast<uint64_t>(tiles),
                 static_cast<uint64_t>(tiles), partial);
    MVS_LAUNCH(s, "curv_mean", k_mean_finish, 1, 1, 0, partial, ne, mean);
    MVS_LAUNCH(s, "curv_edge_tensor", k_edge_tensor, blocks_for(ne, kThreads), kThreads, 0, x, y,
               z, fn, nf, edges, ne, mean, contrib);
  }
  MVS_LAUNCH(s, "curv_greater", k_greater_count, blocks_for(n, kThreads), kThreads, 0, off, idx, n,
             gcnt);
  exclusive_scan_u32(s, gcnt, estart, n);
  MVS_LAUNCH(s, "curv_gather", k_gather, blocks_for(n, kThreads), kThreads, 0, off, idx, estart,
             contrib, ne, n, Ta);
  double *cur = Ta, *nxt = Tb;
  for (int it = 0; it < smoothing; ++it) {
    MVS_LAUNCH(s, "curv_smooth", k_smooth, blocks_for(n, kThreads), kThreads, 0, off, idx, n, cur,
               nxt);
    double *t = cur;
    cur = nxt;
    nxt = t;
  }
  MVS_LAUNCH(s, "curv_frames", k_frames, blocks_for(n, kThreads), kThreads, 0, cur, normals, off,
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
