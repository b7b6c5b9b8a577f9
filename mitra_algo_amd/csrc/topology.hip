// prepare() on the device (src/core/prepare.cpp:13-58 topology, :102-149
// validation and box/centre; subdivision off).  Same outputs as the host
// mvs_prepare, bit for bit, written to device arrays so curvature and the
// stages after it start from HBM.
//
// Layout of the build (F kept faces, 3F corners, E edges, V vertices):
//   * box: per-block min/max partials, one block finishes (min/max do not
//     depend on order, so the box size is the host's exactly);
//   * faces: range check and degenerate test per face, an order-preserving
//     compaction (flags, scan), face normals and areas of the kept faces;
//   * edges: every corner (a, b) emits the key (lo << bits | hi) with its
//     corner id; a stable LSD radix sort orders the keys, so each edge's
//     occurrences stay in face/corner order (the reference's per-edge lists
//     in std::map<(lo, hi)> order).  Group starts become edge records
//     (first occurrence's a, b, face0; the second's face as face1), with
//     the non-manifold and winding counts;
//   * incident faces: corners sorted stably by their vertex: the sorted
//     corner ids divided by 3 are the CSR indices in face order, and the
//     row starts of the sorted keys are its offsets;
//   * rings: row v holds the lo of every edge (lo, v) in lo order (the
//     edges re-sorted stably by hi), then the hi of every edge (v, hi) in hi
//     order (the edge list itself): sorted and unique by construction;
//   * vertex normals and weights: per vertex, its incident faces in face
//     order, summed as the host sums them, then unit();
//   * centre: the vertex mean as a sequential sum in vertex order on the
//     host, while the device builds the rest (a fixed-order parallel sum
//     would differ from it in the last bits).
// Two small read-backs size the sorts (kept faces, then edges).
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <thread>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr unsigned kBoxBlocks = 512;

// Read back after the face pass, then after the edge pass.
struct PrepStatus {
  double box;
  unsigned nonfinite, bad_index, nonmanifold, conflicts;
};

__device__ inline double dmin(double a, double b) { return b < a ? b : a; }
__device__ inline double dmax(double a, double b) { return a < b ? b : a; }

__global__ void __launch_bounds__(kThreads)
    k_box_partial(const double *x, const double *y, const double *z, uint64_t nv, double *part,
                  PrepStatus *st) {
  __shared__ double red[6][kThreads];
  double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
  bool bad = false;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < nv;
       i += static_cast<uint64_t>(gridDim.x) * kThreads) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; ++k) {
      bad |= !isfinite(v[k]);
      lo[k] = dmin(lo[k], v[k]);
      hi[k] = dmax(hi[k], v[k]);
    }
  }
  if (bad)
    atomicOr(&st->nonfinite, 1u);
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = dmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        red[3 + k][threadIdx.x] = dmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6)
    part[threadIdx.x * kBoxBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// box_size = norm(hi - lo) of the per-block partials (prepare.cpp:117-126).
__global__ void __launch_bounds__(kThreads) k_box_finish(const double *part, PrepStatus *st) {
  __shared__ double red[6][kThreads];
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = part[k * kBoxBlocks];
    hi[k] = part[(3 + k) * kBoxBlocks];
  }
  for (unsigned b = threadIdx.x; b < kBoxBlocks; b += kThreads)
    for (int k = 0; k < 3; ++k) {
      lo[k] = dmin(lo[k], part[k * kBoxBlocks + b]);
      hi[k] = dmax(hi[k], part[(3 + k) * kBoxBlocks + b]);
    }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = dmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
        red[3 + k][threadIdx.x] = dmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    st->box = norm(v3(red[3][0], red[4][0], red[5][0]) - v3(red[0][0], red[1][0], red[2][0]));
}

__device__ inline V3 vertex(const double *x, const double *y, const double *z, int32_t i) {
  return v3(x[i], y[i], z[i]);
}

// Range check and degenerate test per input face (prepare.cpp:128-139).
__global__ void k_face_flags(const int32_t *f0, const int32_t *f1, const int32_t *f2, uint64_t nf,
                             uint64_t nv, const double *x, const double *y, const double *z,
                             PrepStatus *st, uint32_t *keep) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nf)
    return;
  const int32_t f[3] = {f0[i], f1[i], f2[i]};
  bool in = true;
  for (int k = 0; k < 3; ++k)
    in &= f[k] >= 0 && static_cast<uint64_t>(f[k]) < nv;
  if (!in) {
    atomicOr(&st->bad_index, 1u);
    keep[i] = 0;
    return;
  }
  const double box = st->box;
  const V3 a = vertex(x, y, z, f[0]), b = vertex(x, y, z, f[1]), c = vertex(x, y, z, f[2]);
  keep[i] = !(norm(cross(b - a, c - a)) <= 1e-14 * box * box);
}

// Kept faces in input order; face normal and area as topology() computes
// them (prepare.cpp:21-25); corner keys for the edge and incidence sorts.
__global__ void k_face_compact(const int32_t *f0, const int32_t *f1, const int32_t *f2,
                               uint64_t nf, const uint32_t *keep, const uint64_t *foff,
                               const double *x, const double *y, const double *z, int bits,
                               int32_t *o0, int32_t *o1, int32_t *o2, double *fnx, double *fny,
                               double *fnz, double *area, uint64_t *ekey, uint32_t *eval,
                               uint64_t *ikey, uint32_t *ival) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nf || !keep[i])
    return;
  const uint64_t j = foff[i];
  const int32_t f[3] = {f0[i], f1[i], f2[i]};
  o0[j] = f[0];
  o1[j] = f[1];
  o2[j] = f[2];
  const V3 p0 = vertex(x, y, z, f[0]), p1 = vertex(x, y, z, f[1]), p2 = vertex(x, y, z, f[2]);
  V3 n = cross(p1 - p0, p2 - p0);
  area[j] = norm(n) / 2;
  n = unit(n);
  fnx[j] = n.x;
  fny[j] = n.y;
  fnz[j] = n.z;
  for (int k = 0; k < 3; ++k) {
    const uint32_t a = static_cast<uint32_t>(f[k]), b = static_cast<uint32_t>(f[(k + 1) % 3]);
    const uint64_t c = 3 * j + k;
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    ekey[c] = (static_cast<uint64_t>(lo) << bits) | hi;
    eval[c] = static_cast<uint32_t>(c);
    ikey[c] = a;
    ival[c] = static_cast<uint32_t>(c);
  }
}

__global__ void k_group_starts(const uint64_t *key, uint64_t m, uint32_t *start) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < m)
    start[j] = j == 0 || key[j] != key[j - 1];
}

// One record per (lo, hi) group (prepare.cpp:43-51): the first occurrence's
// directed corner and face, the second occurrence's face for a two-face
// edge; more than two occurrences count as non-manifold, two with the same
// direction as a winding conflict.  Writes the transposed (hi, lo) entries
// and each edge's lo for the ring build.
__global__ void k_edge_records(const uint64_t *key, const uint32_t *corner, uint64_t m,
                               const uint32_t *start, const uint64_t *eoff, const int32_t *o0,
                               const int32_t *o1, const int32_t *o2, int bits, mvs_edge *edges,
                               uint64_t *tkey, uint32_t *tval, uint32_t *elo, PrepStatus *st) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= m || !start[j])
    return;
  const uint64_t k = key[j];
  const bool two = j + 1 < m && key[j + 1] == k;
  const bool more = two && j + 2 < m && key[j + 2] == k;
  const int32_t *F[3] = {o0, o1, o2};
  const uint32_t c0 = corner[j];
  const int32_t face0 = static_cast<int32_t>(c0 / 3), k0 = static_cast<int32_t>(c0 % 3);
  const int32_t a = F[k0][face0], b = F[(k0 + 1) % 3][face0];
  int32_t face1 = -1;
  if (more)
    atomicAdd(&st->nonmanifold, 1u);
  if (two && !more) {
    const uint32_t c1 = corner[j + 1];
    face1 = static_cast<int32_t>(c1 / 3);
    const int32_t k1 = static_cast<int32_t>(c1 % 3);
    if (a != F[(k1 + 1) % 3][face1])
      atomicAdd(&st->conflicts, 1u);
  }
  const uint64_t e = eoff[j];
  const uint32_t lo = static_cast<uint32_t>(k >> bits);
  const uint32_t hi = static_cast<uint32_t>(k & ((uint64_t(1) << bits) - 1));
  edges[e] = mvs_edge{static_cast<int32_t>(lo), static_cast<int32_t>(hi), a, b, face0, face1};
  tkey[e] = hi;
  tval[e] = lo;
  elo[e] = lo;
}

// start[v] = first position whose row is >= v, for v in [0, rows], over a
// list sorted by row (each v written once).
template <class K>
__global__ void k_row_starts(const K *row, uint64_t m, uint64_t rows, uint64_t *start) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j > m)
    return;
  const uint64_t prev = j == 0 ? 0 : static_cast<uint64_t>(row[j - 1]) + 1;
  const uint64_t cur = j == m ? rows : static_cast<uint64_t>(row[j]);
  for (uint64_t v = prev; v <= cur && v <= rows; ++v)
    start[v] = j;
}

__global__ void k_corner_faces(const uint32_t *corner, uint64_t m, int32_t *idx) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j < m)
    idx[j] = static_cast<int32_t>(corner[j] / 3);
}

__global__ void k_ring_offsets(const uint64_t *lower, const uint64_t *upper, uint64_t nv,
                               uint64_t *off) {
  const uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v <= nv)
    off[v] = lower[v] + upper[v];
}

// Ring row v: lo of every edge (lo, v) in lo order, then hi of every edge
// (v, hi) in hi order.
__global__ void k_ring_fill(const uint64_t *tkey, const uint32_t *tval, const mvs_edge *edges,
                            const uint32_t *elo, uint64_t ne, const uint64_t *lower,
                            const uint64_t *upper, const uint64_t *off, int32_t *idx) {
  const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= ne)
    return;
  const uint64_t v = tkey[j];
  idx[off[v] + (j - lower[v])] = static_cast<int32_t>(tval[j]);
  const uint64_t u = elo[j];
  idx[off[u] + (lower[u + 1] - lower[u]) + (j - upper[u])] = edges[j].hi;
}

// Vertex normal and weight: incident faces in face order, summed as
// prepare.cpp:26-29 sums them, then unit (prepare.cpp:37).
__global__ void k_vertex_normals(const uint64_t *inc_off, const int32_t *inc_idx, uint64_t nv,
                                 const double *fnx, const double *fny, const double *fnz,
                                 const double *area, double *nx, double *ny, double *nz,
                                 double *w) {
  const uint64_t v = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nv)
    return;
  V3 acc = v3(0, 0, 0);
  double wt = 0;
  for (uint64_t j = inc_off[v]; j < inc_off[v + 1]; ++j) {
    const int32_t i = inc_idx[j];
    acc = acc + v3(fnx[i], fny[i], fnz[i]);
    wt += area[i] / 3;
  }
  acc = unit(acc);
  nx[v] = acc.x;
  ny[v] = acc.y;
  nz[v] = acc.z;
  w[v] = wt;
}

int index_bits(uint64_t nv) {
  int bits = 1;
  while (bits < 32 && (uint64_t(1) << bits) < nv)
    ++bits;
  return bits;
}

}  // namespace

void prepare_device(Session *s, int64_t nv_, const double *hx, const double *hy, const double *hz,
                    int64_t nf_, const int32_t *hf0, const int32_t *hf1, const int32_t *hf2,
                    double *xyz, mvs_prepared *o) {
  if (!o || !xyz || !hx || !hy || !hz || (nf_ > 0 && (!hf0 || !hf1 || !hf2)))
    raise(MVS_EINVAL, "null argument");
  if (nv_ <= 0 || nf_ <= 0)
    raise(MVS_ERUNTIME, "Mesh is empty");
  if (nv_ > std::numeric_limits<int32_t>::max() || nf_ > std::numeric_limits<int32_t>::max())
    raise(MVS_ERUNTIME, "Mesh exceeds 32-bit index capacity");
  const uint64_t nv = static_cast<uint64_t>(nv_), nf = static_cast<uint64_t>(nf_);
  const int bits = index_bits(nv);
  double *x = xyz, *y = xyz + nv, *z = xyz + 2 * nv;
  // the centre on a host thread while the uploads and the device build run
  // (prepare.cpp:145-148)
  V3 c = v3(0, 0, 0);
  std::thread centre([&] {
    V3 acc = v3(0, 0, 0);
    for (uint64_t i = 0; i < nv; ++i)
      acc = acc + v3(hx[i], hy[i], hz[i]);
    c = vdivs(acc, static_cast<double>(nv));
  });
  struct Join {
    std::thread &t;
    ~Join() {
      if (t.joinable())  // an error left early
        t.join();
    }
  } join{centre};
  int32_t *fin = scratch_as<int32_t>(s, kSlotPrep0, 3 * nf);
  copy_h2d(s, x, hx, nv * 8);
  copy_h2d(s, y, hy, nv * 8);
  copy_h2d(s, z, hz, nv * 8);
  copy_h2d(s, fin, hf0, nf * 4);
  copy_h2d(s, fin + nf, hf1, nf * 4);
  copy_h2d(s, fin + 2 * nf, hf2, nf * 4);

  PrepStatus *st = scratch_as<PrepStatus>(s, kSlotPrep1, 1);
  double *part = scratch_as<double>(s, kSlotPrep2, 6 * kBoxBlocks);
  uint32_t *keep = scratch_as<uint32_t>(s, kSlotPrep3, nf);
  uint64_t *foff = scratch_as<uint64_t>(s, kSlotPrep4, nf + 1);
  zero_device(s, st, sizeof(PrepStatus));
  const unsigned vb = blocks_for(nv, kThreads);
  MVS_LAUNCH(s, "prep_box", k_box_partial, kBoxBlocks, kThreads, 0, x, y, z, nv, part, st);
  MVS_LAUNCH(s, "prep_box_finish", k_box_finish, 1, kThreads, 0, part, st);
  MVS_LAUNCH(s, "prep_face_flags", k_face_flags, blocks_for(nf, kThreads), kThreads, 0, fin,
             fin + nf, fin + 2 * nf, nf, nv, x, y, z, st, keep);
  exclusive_scan_u32(s, keep, foff, nf);

  PrepStatus h;
  copy_d2h(s, &h, st, sizeof h);
  uint64_t F = 0;
  copy_d2h(s, &F, foff + nf, sizeof F);
  if (h.nonfinite)
    raise(MVS_ERUNTIME, "Non-finite coordinate");
  if (h.box < 1e-14)
    raise(MVS_ERUNTIME, "Mesh has zero extent");
  if (h.bad_index)
    raise(MVS_ERUNTIME, "Face index out of range");
  if (!F)
    raise(MVS_ERUNTIME, "No nondegenerate triangles");
  if (F > std::numeric_limits<uint32_t>::max() / 3)
    raise(MVS_EOVERFLOW, "the device topology build takes at most 2^32 / 3 faces");
  const uint64_t M = 3 * F;
  double *area = scratch_as<double>(s, kSlotPrep5, F);
  uint64_t *ekey = scratch_as<uint64_t>(s, kSlotPrep6, M);
  uint32_t *eval = scratch_as<uint32_t>(s, kSlotPrep7, M);
  uint64_t *ikey = scratch_as<uint64_t>(s, kSlotPrep8, M);
  uint32_t *ival = scratch_as<uint32_t>(s, kSlotPrep9, M);
  MVS_LAUNCH(s, "prep_face_compact", k_face_compact, blocks_for(nf, kThreads), kThreads, 0, fin,
             fin + nf, fin + 2 * nf, nf, keep, foff, x, y, z, bits, o->f0, o->f1, o->f2, o->fnx,
             o->fny, o->fnz, area, ekey, eval, ikey, ival);
  uint64_t *ekey2 = scratch_as<uint64_t>(s, kSlotPrep10, M);
  uint32_t *eval2 = scratch_as<uint32_t>(s, kSlotPrep11, M);
  radix_sort_pairs_u64(s, ekey, eval, ekey2, eval2, M, 2 * bits);
  uint32_t *gstart = scratch_as<uint32_t>(s, kSlotPrep17, M);
  uint64_t *eoff = scratch_as<uint64_t>(s, kSlotPrep18, M + 1);
  MVS_LAUNCH(s, "prep_edge_starts", k_group_starts, blocks_for(M, kThreads), kThreads, 0, ekey2,
             M, gstart);
  exclusive_scan_u32(s, gstart, eoff, M);
  // edge-indexed arrays (E <= M)
  uint64_t *tkey = ekey, *tkey2 = scratch_as<uint64_t>(s, kSlotPrep12, M);
  uint32_t *tval = eval, *tval2 = scratch_as<uint32_t>(s, kSlotPrep13, M);
  uint32_t *elo = scratch_as<uint32_t>(s, kSlotPrep14, M);
  MVS_LAUNCH(s, "prep_edge_records", k_edge_records, blocks_for(M, kThreads), kThreads, 0, ekey2,
             eval2, M, gstart, eoff, o->f0, o->f1, o->f2, bits, o->edges, tkey, tval, elo, st);
  // incident faces: corners by vertex, stable
  radix_sort_pairs_u64(s, ikey, ival, ekey2, eval2, M, bits);
  MVS_LAUNCH(s, "prep_inc_offsets", k_row_starts<uint64_t>, blocks_for(M + 1, kThreads), kThreads,
             0, ekey2, M, nv, o->inc_off);
  MVS_LAUNCH(s, "prep_inc_faces", k_corner_faces, blocks_for(M, kThreads), kThreads, 0, eval2, M,
             o->inc_idx);
  MVS_LAUNCH(s, "prep_vertex_normals", k_vertex_normals, vb, kThreads, 0, o->inc_off, o->inc_idx,
             nv, o->fnx, o->fny, o->fnz, area, o->nx, o->ny, o->nz, o->weights);

  uint64_t E = 0;
  copy_d2h(s, &E, eoff + M, sizeof E);
  copy_d2h(s, &h, st, sizeof h);
  if (h.nonmanifold || h.conflicts)
    raise(MVS_ERUNTIME, "Mesh must have manifold edges and consistent face winding; repair "
                        "input mesh first");
  // rings: the edges re-sorted by hi (stable: lo order within a row)
  radix_sort_pairs_u64(s, tkey, tval, tkey2, tval2, E, bits);
  uint64_t *lower = scratch_as<uint64_t>(s, kSlotPrep15, nv + 1);
  uint64_t *upper = scratch_as<uint64_t>(s, kSlotPrep16, nv + 1);
  MVS_LAUNCH(s, "prep_ring_lower", k_row_starts<uint64_t>, blocks_for(E + 1, kThreads), kThreads,
             0, tkey2, E, nv, lower);
  MVS_LAUNCH(s, "prep_ring_upper", k_row_starts<uint32_t>, blocks_for(E + 1, kThreads), kThreads,
             0, elo, E, nv, upper);
  MVS_LAUNCH(s, "prep_ring_offsets", k_ring_offsets, blocks_for(nv + 1, kThreads), kThreads, 0,
             lower, upper, nv, o->nbr_off);
  MVS_LAUNCH(s, "prep_ring_fill", k_ring_fill, blocks_for(E, kThreads), kThreads, 0, tkey2, tval2,
             o->edges, elo, E, lower, upper, o->nbr_off, o->nbr_idx);

  o->face_count = static_cast<int64_t>(F);
  o->nbr_nnz = static_cast<int64_t>(2 * E);
  o->inc_nnz = static_cast<int64_t>(M);
  o->edge_count = static_cast<int64_t>(E);
  o->box_size = h.box;
  centre.join();
  o->center[0] = c.x;
  o->center[1] = c.y;
  o->center[2] = c.z;
  o->rejected_faces = static_cast<int32_t>(nf - F);
  o->nonmanifold_edges = 0;
  o->orientation_conflicts = 0;
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_prepare(mvs_session *sp, int64_t nv, const double *x, const double *y,
                    const double *z, int64_t nf, const int32_t *f0, const int32_t *f1,
                    const int32_t *f2, double *xyz, mvs_prepared *out) {
  return guarded([&] {
    Session *s = as_session(sp);
    prepare_device(s, nv, x, y, z, nf, f0, f1, f2, xyz, out);
  });
}

}  // extern "C"
