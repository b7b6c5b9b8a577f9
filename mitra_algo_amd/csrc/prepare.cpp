// Host mesh preparation and topology (restates src/core/prepare.cpp:13-58
// and :102-149 with subdivision off).
//
// The reference builds edges in a std::map keyed by (lo, hi) with one node
// allocation per edge.  Here every face corner emits a 64-bit key
// (lo << 32 | hi); a stable LSD radix sort over the keys yields the same
// (lo, hi) order with each edge's face occurrences in face/corner order,
// which is what the reference's per-edge vectors hold.  Normals and weights
// accumulate in face order, rings are sorted unique per vertex, incident
// faces are listed in face order, exactly as in the reference, so curvature
// sums downstream see the same operand order.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "common.hpp"

namespace mvs {
namespace {

void radix_sort_keys(std::vector<uint64_t> &keys, std::vector<uint32_t> &idx) {
  const size_t n = keys.size();
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> i2(n);
  std::vector<size_t> count(1 << 16);
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 16 * pass;
    std::fill(count.begin(), count.end(), 0);
    for (size_t i = 0; i < n; ++i)
      ++count[(keys[i] >> shift) & 0xFFFF];
    size_t run = 0;
    for (auto &c : count) {
      size_t t = c;
      c = run;
      run += t;
    }
    for (size_t i = 0; i < n; ++i) {
      size_t d = (keys[i] >> shift) & 0xFFFF;
      size_t p = count[d]++;
      k2[p] = keys[i];
      i2[p] = idx[i];
    }
    keys.swap(k2);
    idx.swap(i2);
  }
}

}  // namespace
}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_prepare(int64_t nv, const double *x, const double *y, const double *z, int64_t nf,
                const int32_t *f0, const int32_t *f1, const int32_t *f2, mvs_prepared *o) {
  return guarded([&] {
    if (!o)
      raise(MVS_EINVAL, "null output");
    if (nv <= 0 || nf <= 0)
      raise(MVS_ERUNTIME, "Mesh is empty");
    if (nv > std::numeric_limits<int32_t>::max() || nf > std::numeric_limits<int32_t>::max())
      raise(MVS_ERUNTIME, "Mesh exceeds 32-bit index capacity");
    double lo[3] = {x[0], y[0], z[0]}, hi[3] = {x[0], y[0], z[0]};
    for (int64_t i = 0; i < nv; ++i) {
      double v[3] = {x[i], y[i], z[i]};
      for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(v[k]))
          raise(MVS_ERUNTIME, "Non-finite coordinate");
        lo[k] = v[k] < lo[k] ? v[k] : lo[k];
        hi[k] = hi[k] < v[k] ? v[k] : hi[k];
      }
    }
    o->box_size = norm(v3(hi[0], hi[1], hi[2]) - v3(lo[0], lo[1], lo[2]));
    if (o->box_size < 1e-14)
      raise(MVS_ERUNTIME, "Mesh has zero extent");
    o->rejected_faces = o->nonmanifold_edges = o->orientation_conflicts = 0;
    int64_t F = 0;
    for (int64_t i = 0; i < nf; ++i) {
      int32_t f[3] = {f0[i], f1[i], f2[i]};
      for (int k = 0; k < 3; ++k)
        if (f[k] < 0 || f[k] >= nv)
          raise(MVS_ERUNTIME, "Face index out of range");
      V3 a = v3(x[f[0]], y[f[0]], z[f[0]]), b = v3(x[f[1]], y[f[1]], z[f[1]]),
         c = v3(x[f[2]], y[f[2]], z[f[2]]);
      if (norm(cross(b - a, c - a)) <= 1e-14 * o->box_size * o->box_size) {
        ++o->rejected_faces;
        continue;
      }
      o->f0[F] = f[0], o->f1[F] = f[1], o->f2[F] = f[2];
      ++F;
    }
    o->face_count = F;
    if (!F)
      raise(MVS_ERUNTIME, "No nondegenerate triangles");
    for (int64_t i = 0; i < nv; ++i)
      o->nx[i] = o->ny[i] = o->nz[i] = o->weights[i] = 0;
    std::vector<uint32_t> ring_count(nv, 0), inc_count(nv, 0);
    std::vector<uint64_t> keys(3 * F);
    std::vector<uint32_t> order(3 * F);
    for (int64_t i = 0; i < F; ++i) {
      int32_t f[3] = {o->f0[i], o->f1[i], o->f2[i]};
      V3 p0 = v3(x[f[0]], y[f[0]], z[f[0]]), p1 = v3(x[f[1]], y[f[1]], z[f[1]]),
         p2 = v3(x[f[2]], y[f[2]], z[f[2]]);
      V3 n = cross(p1 - p0, p2 - p0);
      double area = norm(n) / 2;
      n = unit(n);
      o->fnx[i] = n.x, o->fny[i] = n.y, o->fnz[i] = n.z;
      for (int k = 0; k < 3; ++k) {
        int32_t a = f[k], b = f[(k + 1) % 3];
        V3 acc = v3(o->nx[a], o->ny[a], o->nz[a]) + n;
        o->nx[a] = acc.x, o->ny[a] = acc.y, o->nz[a] = acc.z;
        o->weights[a] += area / 3;
        ++ring_count[a];
        ++ring_count[b];
        ++inc_count[a];
        uint32_t l = static_cast<uint32_t>(a < b ? a : b), h = static_cast<uint32_t>(a < b ? b : a);
        keys[3 * i + k] = (static_cast<uint64_t>(l) << 32) | h;
        order[3 * i + k] = static_cast<uint32_t>(3 * i + k);
      }
    }
    // rings: rows with duplicates, then sorted unique
    std::vector<uint64_t> roff(nv + 1, 0);
    for (int64_t i = 0; i < nv; ++i)
      roff[i + 1] = roff[i] + ring_count[i];
    std::vector<int32_t> rbuf(roff[nv]);
    std::vector<uint32_t> rfill(nv, 0), ifill(nv, 0);
    o->inc_off[0] = 0;
    for (int64_t i = 0; i < nv; ++i)
      o->inc_off[i + 1] = o->inc_off[i] + inc_count[i];
    for (int64_t i = 0; i < F; ++i) {
      int32_t f[3] = {o->f0[i], o->f1[i], o->f2[i]};
      for (int k = 0; k < 3; ++k) {
        int32_t a = f[k], b = f[(k + 1) % 3];
        rbuf[roff[a] + rfill[a]++] = b;
        rbuf[roff[b] + rfill[b]++] = a;
        o->inc_idx[o->inc_off[a] + ifill[a]++] = static_cast<int32_t>(i);
      }
    }
    o->inc_nnz = static_cast<int64_t>(o->inc_off[nv]);
    int64_t nnz = 0;
    o->nbr_off[0] = 0;
    for (int64_t i = 0; i < nv; ++i) {
      V3 acc = unit(v3(o->nx[i], o->ny[i], o->nz[i]));
      o->nx[i] = acc.x, o->ny[i] = acc.y, o->nz[i] = acc.z;
      int32_t *row = rbuf.data() + roff[i];
      int64_t len = static_cast<int64_t>(ring_count[i]);
      for (int64_t a = 1; a < len; ++a) {  // rows are short: insertion sort
        int32_t v = row[a];
        int64_t b = a - 1;
        while (b >= 0 && row[b] > v) {
          row[b + 1] = row[b];
          --b;
        }
        row[b + 1] = v;
      }
      for (int64_t k = 0; k < len; ++k)
        if (k == 0 || row[k] != row[k - 1])
          o->nbr_idx[nnz++] = row[k];
      o->nbr_off[i + 1] = static_cast<uint64_t>(nnz);
    }
    o->nbr_nnz = nnz;
    // edges in (lo, hi) order; occurrences of one edge stay in face/corner order
    radix_sort_keys(keys, order);
    int64_t E = 0;
    const size_t M = keys.size();
    for (size_t s = 0; s < M;) {
      size_t t = s + 1;
      while (t < M && keys[t] == keys[s])
        ++t;
      size_t size = t - s;
      uint32_t c0 = order[s];
      int32_t face0 = static_cast<int32_t>(c0 / 3), k0 = static_cast<int32_t>(c0 % 3);
      int32_t fa[3] = {o->f0[face0], o->f1[face0], o->f2[face0]};
      int32_t a = fa[k0], b = fa[(k0 + 1) % 3];
      int32_t face1 = -1;
      if (size > 2)
        ++o->nonmanifold_edges;
      if (size == 2) {
        uint32_t c1 = order[s + 1];
        face1 = static_cast<int32_t>(c1 / 3);
        int32_t k1 = static_cast<int32_t>(c1 % 3);
        int32_t fb[3] = {o->f0[face1], o->f1[face1], o->f2[face1]};
        if (a != fb[(k1 + 1) % 3])
          ++o->orientation_conflicts;
      }
      mvs_edge e = {static_cast<int32_t>(keys[s] >> 32), static_cast<int32_t>(keys[s] & 0xFFFFFFFF),
                    a, b, face0, face1};
      o->edges[E++] = e;
      s = t;
    }
    o->edge_count = E;
    if (o->nonmanifold_edges || o->orientation_conflicts)
      raise(MVS_ERUNTIME, "Mesh must have manifold edges and consistent face winding; repair "
                          "input mesh first");
    V3 c = v3(0, 0, 0);
    for (int64_t i = 0; i < nv; ++i)
      c = c + v3(x[i], y[i], z[i]);
    c = vdivs(c, static_cast<double>(nv));
    o->center[0] = c.x, o->center[1] = c.y, o->center[2] = c.z;
  });
}

}  // extern "C"
