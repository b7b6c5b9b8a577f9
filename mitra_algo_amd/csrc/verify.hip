// Verification of detected transforms against the mesh surface
// (Backend::verification, include/symmetry/pipeline.h:67-68;
// src/cpu/verification.cpp:50-110) and its patches (connected components of
// the accepted vertices over the one-ring graph, verification.cpp:22-49).
//
// For every cluster c and vertex i, two rows: side A maps p_i forward
// (q = R p + t), side B backward (q = R^T (p - t)).  A row is accepted when
//   |q - p| >= 2 tau, the reflection side test passes (A: n.p + d < -tau,
//   B: n.p + d > tau; rigid clusters skip it), the nearest mesh face to q is
//   within tau (squared distance <= tau^2) and the transformed vertex normal
//   meets that face's unit normal at dot >= 0.5,
// with tau = verify_threshold * box.  rms = sqrt(sum of side-A squared
// distances / |A|), coverage = min(1, area(A) + area(B) over the total area).
//
// Exactness.  Every FP64 expression keeps the reference's operation order
// (this library builds with -ffp-contract=off).  The nearest face comes from
// a tree built exactly like the reference's surface tree
// (src/cpu/surface.h:16-48): boxes over face corners, split at the median
// face (std::nth_element by centroid along the widest axis) while a node holds
// more than 8 faces, left half first; and it is searched in the same order
// (surface.h:57-81: skip a node whose box is farther than the best distance,
// scan a leaf's faces in order keeping strictly smaller distances, visit the
// nearer child first).  The minimum distance is therefore the same value, and
// among faces at exactly that distance the same face wins, so the normal test
// and with it the accepted sets are the reference's.
//
// Device work: one thread per row (k_verify, FP64, per-thread stack in LDS),
// a union-find over the accepted vertices whose roots are the smallest member
// (k_cc_init / k_cc_hook / k_cc_flatten), and an ordered compaction of the
// accepted rows.
// The host groups the compact lists into patches in O(|side|).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

#include "session.hpp"

namespace mvs {
namespace {

constexpr int kVThreads = 256;
constexpr int kLeafFaces = 8;
constexpr int kFaceWords = 10;  // a, b, c, degenerate flag (normals apart)
constexpr int kJumpRounds = 10;  // pointer-jumping rounds before the hooks

// Tree node: box, children (left < 0 for a leaf) and the leaf's face range
// in leaf order.
struct alignas(16) SNode {
  double lo[3], hi[3];
  int32_t left, right, begin, end;
};
static_assert(sizeof(SNode) == 64, "node layout");

// Search record of an inner node: its two children's boxes and references
// (one 128-byte line per visited inner node).  A reference >= 0 is an inner
// node; a leaf of n faces starting at leaf-ordered face b is -1 - (16 b + n).
struct alignas(128) Inner {
  double lo[2][3], hi[2][3];
  int32_t ref[2];
};
static_assert(sizeof(Inner) == 128, "search record layout");
constexpr uint64_t kMaxLeafStart = uint64_t{1} << 27;  // 16 b + n fits int32

// One cluster's motion (ClusterRecord, include/symmetry/storage.h:365-370).
struct Motion {
  double r[9], t[3], n[3], off;
  int32_t reflection, pad;
};

// ------------------------------------------------------------ host tree
// Surface tree in the reference's build order (surface.h:16-48), pre-order
// node numbering.  A node of n faces splits into n / 2 faces on the left and
// the rest on the right, so the node count and height of a subtree depend on
// n alone: the right subtree's first node is known before the left one is
// built, and independent subtrees are built by separate threads.  The
// partitions are those of std::nth_element run on the same ranges with the
// same comparisons, whichever thread runs them.
struct TreeBuilder {
  const double *x, *y, *z;
  const int32_t *f[3];
  std::vector<double> cen[3];
  std::vector<double> fbox;  // per face: corner minimum xyz, maximum xyz
  std::vector<int> ids;
  std::vector<SNode> nodes;
  std::unordered_map<int, std::pair<int, int>> shape;  // faces -> (nodes, height)

  double coord(int v, int k) const { return k == 0 ? x[v] : (k == 1 ? y[v] : z[v]); }

  std::pair<int, int> shape_of(int n) {
    auto it = shape.find(n);
    if (it != shape.end())
      return it->second;
    std::pair<int, int> r{1, 0};
    if (n > kLeafFaces) {
      const auto l = shape_of(n / 2), h = shape_of(n - n / 2);
      r = {1 + l.first + h.first, 1 + std::max(l.second, h.second)};
    }
    shape.emplace(n, r);
    return r;
  }

  void build(int index, int begin, int end, int spawn) {
    SNode n;
    for (int k = 0; k < 3; ++k)
      n.lo[k] = 1e300, n.hi[k] = -1e300;
    // the minimum over the faces' corner minima is the minimum over corners
    for (int i = begin; i < end; ++i) {
      const double *b = fbox.data() + 6 * static_cast<size_t>(ids[i]);
      for (int k = 0; k < 3; ++k) {
        n.lo[k] = std::min(n.lo[k], b[k]);
        n.hi[k] = std::max(n.hi[k], b[3 + k]);
      }
    }
    n.left = n.right = -1;
    n.begin = begin;
    n.end = end;
    if (end - begin > kLeafFaces) {
      const double ex = n.hi[0] - n.lo[0], ey = n.hi[1] - n.lo[1], ez = n.hi[2] - n.lo[2];
      int axis = ey > ex ? 1 : 0;
      if (ez > (axis == 1 ? ey : ex))
        axis = 2;
      const int mid = (begin + end) / 2;
      const double *c = cen[axis].data();
      std::nth_element(ids.begin() + begin, ids.begin() + mid, ids.begin() + end,
                       [c](int a, int b) { return c[a] < c[b]; });
      n.left = index + 1;
      n.right = index + 1 + shape.at(mid - begin).first;
      nodes[index] = n;
      if (spawn > 0 && end - begin > (1 << 15)) {
        std::thread t([&, n, mid, end, spawn] { build(n.right, mid, end, spawn - 1); });
        build(n.left, begin, mid, spawn - 1);
        t.join();
      } else {
        build(n.left, begin, mid, 0);
        build(n.right, mid, end, 0);
      }
      return;
    }
    nodes[index] = n;
  }
};

// fn(first, last) over [0, n) in contiguous chunks on up to 16 threads.
template <class F> void parallel_chunks(uint64_t n, F fn) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const uint64_t parts = std::min<uint64_t>(hw, (n + 65535) / 65536);
  if (parts <= 1) {
    fn(uint64_t{0}, n);
    return;
  }
  std::vector<std::thread> ts;
  for (uint64_t t = 1; t < parts; ++t)
    ts.emplace_back([=] { fn(n * t / parts, n * (t + 1) / parts); });
  fn(uint64_t{0}, n / parts);
  for (auto &t : ts)
    t.join();
}

// -------------------------------------------------------- device geometry
// std::clamp and std::min as the reference evaluates them.
__device__ inline double clamp01(double v) { return v < 0.0 ? 0.0 : (1.0 < v ? 1.0 : v); }
__device__ inline double min2(double a, double b) { return b < a ? b : a; }

// Squared distance from p to a box (surface.h:49-56).  The reference takes
// std::max({lo - p, 0, p - hi}); fmax gives the same value up to the sign of
// a zero, which the square discards.
__device__ inline double box_bound(V3 p, const double *lo, const double *hi) {
  double s = 0;
  double d = fmax(fmax(lo[0] - p.x, 0.0), p.x - hi[0]);
  s += d * d;
  d = fmax(fmax(lo[1] - p.y, 0.0), p.y - hi[1]);
  s += d * d;
  d = fmax(fmax(lo[2] - p.z, 0.0), p.z - hi[2]);
  s += d * d;
  return s;
}

// Squared distance from p to segment ab (src/core/math.cpp:65-71).
__device__ inline double seg_d2(V3 p, V3 a, V3 b) {
  V3 d = b - a;
  double l = dot(d, d);
  double t = l > 1e-14 ? clamp01(dot(p - a, d) / l) : 0.0;
  V3 e = p - (a + d * t);
  return dot(e, e);
}

// Squared distance from p to triangle abc by the Voronoi regions of the
// vertices, edges and face (Ericson, Real-Time Collision Detection, 2005,
// 5.1.5), with the reference's degenerate fallback and operation order
// (src/core/math.cpp:72-112).  `degenerate` is the reference's test
// |ab x ac|^2 < 1e-28, which does not depend on p and is evaluated once per
// face on the host with the same operations.
//
// Without branches: every dot product of the reference is formed up front
// (forming one early changes no value), the region is the first whose test
// holds, in the reference's order, and each region's closest point is
// written in one form, base + d1 s + d2 t:
//   vertex a, b, c  base = a, b, c            s = t = 0
//   edge ab         base = a, d1 = ab,        s = d1 / (d1 - d3)
//   edge ac         base = a, d1 = ac,        s = d2 / (d2 - d6)
//   edge bc         base = b, d1 = c - b,     s = (d4-d3) / ((d4-d3) + (d5-d6))
//   face            base = a, d1 = ab, d2 = ac,
//                   s = vb / (va + vb + vc),  t = vc / (va + vb + vc)
// Adding a zero product only flips the sign of a zero coordinate of
// e = p - point, which its square discards, so every region's squared
// distance is the reference's to the bit.  The lanes of a wave then run one
// instruction stream instead of every branch in turn.  ops += the FP64
// additions, subtractions, multiplications and divisions of the reference's
// path for that region (24 to 82; 102 for a degenerate triangle).
__device__ inline double tri_d2(V3 p, V3 a, V3 b, V3 c, bool degenerate, uint32_t &ops) {
  if (degenerate) {
    ops += 102;
    return min2(min2(seg_d2(p, a, b), seg_d2(p, a, c)), seg_d2(p, b, c));
  }
  const V3 ab = b - a, ac = c - a, ap = p - a, bp = p - b, cp = p - c;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp),
               d5 = dot(ab, cp), d6 = dot(ac, cp);
  const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
  const double e43 = d4 - d3, e56 = d5 - d6;
  const bool ra = d1 <= 0 && d2 <= 0;
  const bool rb = !ra && d3 >= 0 && d4 <= d3;
  const bool rab = !ra && !rb && vc <= 0 && d1 >= 0 && d3 <= 0;
  const bool rc = !ra && !rb && !rab && d6 >= 0 && d5 <= d6;
  const bool rac = !ra && !rb && !rab && !rc && vb <= 0 && d2 >= 0 && d6 <= 0;
  const bool rbc = !ra && !rb && !rab && !rc && !rac && va <= 0 && e43 >= 0 && e56 >= 0;
  const bool face = !(ra || rb || rab || rc || rac || rbc);
  ops += ra ? 24 : rb ? 37 : rab ? 51 : rc ? 53 : rac ? 67 : rbc ? 75 : 82;
  const V3 base = (rb || rbc) ? b : (rc ? c : a);
  const V3 d1v = rac ? ac : (rbc ? c - b : ab);
  const double num = rab ? d1 : (rac ? d2 : (rbc ? e43 : vb));
  const double den = rab ? d1 - d3 : (rac ? d2 - d6 : (rbc ? e43 + e56 : va + vb + vc));
  const bool vertex = ra || rb || rc;
  const double sv = vertex ? 0.0 : num / den;
  const double tv = face ? vc / (va + vb + vc) : 0.0;
  const V3 e = p - (base + d1v * sv + ac * tv);
  const double r = dot(e, e);
  return 0.0 < r ? r : 0.0;
}

__device__ inline V3 face_corner(const double *fr, int k) {
  return v3(fr[3 * k], fr[3 * k + 1], fr[3 * k + 2]);
}

// ------------------------------------------------------------- kernels
// Row r = (2c + s) * nv + i.  accepted[r] in {0, 1}; label[r] = i when
// accepted, else -1 (the union-find parent array); dist2[r] = the nearest
// squared distance when the row reached the surface search, else NaN.
// One row: transform, early tests, tree search, normal test; tests, visits
// and ops count the triangle tests, the stack entries taken and the FP64
// operations of the search (box and triangle distances).
__device__ inline void verify_row(const double *__restrict__ vx, const double *__restrict__ vy,
                                  const double *__restrict__ vz, const double *__restrict__ nx,
                                  const double *__restrict__ ny, const double *__restrict__ nz,
                                  uint32_t nv, uint64_t row, const uint32_t *__restrict__ order,
                                  const Inner *__restrict__ inner,
                                  int root, const double *__restrict__ faces,
                                  const double *__restrict__ normals,
                                  const Motion *__restrict__ motions, double threshold,
                                  int *stk, float *bnd, uint32_t *__restrict__ accepted,
                                  int32_t *__restrict__ label, double *__restrict__ dist2,
                                  uint32_t &tests, uint32_t &visits, uint32_t &ops) {
  const uint64_t seg = row / nv;
  const uint32_t i = order[row - seg * nv];
  row = seg * nv + i;
  const Motion &mo = motions[seg >> 1];
  const bool inverse = seg & 1;
  uint32_t ok = 0;
  double best = __builtin_nan("");
  const V3 p = v3(vx[i], vy[i], vz[i]);
  M3 r;
  for (int k = 0; k < 9; ++k)
    r.m[k] = mo.r[k];
  if (inverse) {
    // transpose(c.rotation) (verification.cpp:59)
    M3 t = {{r.m[0], r.m[3], r.m[6], r.m[1], r.m[4], r.m[7], r.m[2], r.m[5], r.m[8]}};
    r = t;
  }
  const V3 tr = v3(mo.t[0], mo.t[1], mo.t[2]);
  const V3 q = inverse ? apply(r, p - tr) : apply(r, p) + tr;
  bool go = !(norm(q - p) < threshold * 2);
  if (go && mo.reflection) {
    const double side = dot(v3(mo.n[0], mo.n[1], mo.n[2]), p) + mo.off;
    go = inverse ? !(side <= threshold) : !(side >= -threshold);
  }
  if (go) {
    // Depth-first, nearer child first (the reference's order, decided on
    // exact FP64 box distances).  A stack entry carries its box distance
    // rounded down to FP32: an entry is skipped only when that lower bound
    // already exceeds the best distance, so no subtree holding a face at the
    // minimum is skipped; an entry the reference would skip and this search
    // visits holds only farther faces.  The search alternates between
    // descending to the next leaf and scanning it, so the lanes of a wave do
    // the same kind of work at a time.
    best = __builtin_inf();
    int hit = -1, top = 1;
    stk[0] = root;
    bnd[0] = -__builtin_inff();
    while (true) {
      int leaf = 0;
      while (top > 0) {
        --top;
        const int ref = stk[top * kVThreads];
        const float lb = bnd[top * kVThreads];
        ++visits;
        if (static_cast<double>(lb) > best)
          continue;
        if (ref < 0) {
          leaf = ref;
          break;
        }
        const Inner nd = inner[ref];
        ops += 24;  // two box distances
        const double b0 = box_bound(q, nd.lo[0], nd.hi[0]), b1 = box_bound(q, nd.lo[1], nd.hi[1]);
        const bool swap = b0 > b1;  // the right child first only when strictly nearer
        stk[top * kVThreads] = swap ? nd.ref[0] : nd.ref[1];
        bnd[top * kVThreads] = __double2float_rd(swap ? b0 : b1);
        ++top;
        // at most one waiting sibling per level plus the two children: tree
        // depth + 1 entries suffice
        stk[top * kVThreads] = swap ? nd.ref[1] : nd.ref[0];
        bnd[top * kVThreads] = __double2float_rd(swap ? b1 : b0);
        ++top;
      }
      if (leaf == 0)
        break;
      const int code = -1 - leaf, first = code >> 4, end = first + (code & 15);
      for (int k = first; k < end; ++k) {
        const double *fr = faces + static_cast<uint64_t>(k) * kFaceWords;
        const double d = tri_d2(q, face_corner(fr, 0), face_corner(fr, 1), face_corner(fr, 2),
                                fr[9] != 0.0, ops);
        ++tests;
        if (d < best) {
          best = d;
          hit = k;
        }
      }
    }
    if (hit >= 0 && !(best > threshold * threshold)) {
      const double *fn = normals + 3 * static_cast<uint64_t>(hit);
      const V3 vn = apply(r, v3(nx[i], ny[i], nz[i]));
      ok = dot(vn, v3(fn[0], fn[1], fn[2])) < 0.5 ? 0u : 1u;
    }
  }
  accepted[row] = ok;
  label[row] = ok ? static_cast<int32_t>(i) : -1;
  dist2[row] = best;
}

__global__ void __launch_bounds__(kVThreads)
k_verify(const double *__restrict__ vx, const double *__restrict__ vy,
         const double *__restrict__ vz, const double *__restrict__ nx,
         const double *__restrict__ ny, const double *__restrict__ nz, uint32_t nv,
         uint64_t rows, const Inner *__restrict__ inner, int root,
         const double *__restrict__ faces, const double *__restrict__ normals,
         const Motion *__restrict__ motions, double threshold, int stack_size,
         const uint32_t *__restrict__ order, uint32_t *__restrict__ accepted, int32_t *__restrict__ label,
         double *__restrict__ dist2, unsigned long long *__restrict__ stats) {
  // per thread: stack_size node references, then stack_size FP32 bounds
  extern __shared__ int stack_lds[];
  int *stk = stack_lds + threadIdx.x;
  float *bnd = reinterpret_cast<float *>(stack_lds + stack_size * kVThreads) + threadIdx.x;
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  uint32_t tests = 0, visits = 0, ops = 0;
  if (row < rows)
    verify_row(vx, vy, vz, nx, ny, nz, nv, row, order, inner, root, faces, normals, motions,
               threshold, stk, bnd, accepted, label, dist2, tests, visits, ops);
  // work counters: one atomic per wave
  const uint32_t wt = wave_sum(tests), wv = wave_sum(visits);
  const unsigned long long wo = wave_sum(static_cast<unsigned long long>(ops));
  if (lane_id() == 0) {
    atomicAdd(stats, static_cast<unsigned long long>(wt));
    atomicAdd(stats + 1, static_cast<unsigned long long>(wv));
    atomicAdd(stats + 2, wo);
  }
}

__device__ inline int32_t load_parent(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_parent(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Root of x.  Every parent is smaller than its child (a root is its own
// parent), so the walk ends; on the way each visited vertex is pointed at
// its grandparent, still an ancestor (intermediate pointer jumping).
__device__ inline int32_t cc_find(int32_t *par, int32_t x) {
  int32_t cur = load_parent(par + x);
  if (cur == x)
    return x;
  int32_t prev = x, next;
  while (cur > (next = load_parent(par + cur))) {
    store_parent(par + prev, next);
    prev = cur;
    cur = next;
  }
  return cur;
}

// Union-find over the accepted vertices of each (cluster, side) segment,
// after Jaiganesh and Burtscher, "A high-performance connected components
// implementation for GPUs" (HPDC 2018): every parent is smaller than its
// child, so each root is the smallest member of its component.
// k_cc_init: an accepted vertex starts under its smallest accepted
// neighbour below it (itself when none).
__global__ void __launch_bounds__(kVThreads)
k_cc_init(const uint64_t *__restrict__ nbr_off, const int32_t *__restrict__ nbr_idx,
          const uint32_t *__restrict__ accepted, uint32_t nv, uint64_t rows, int32_t *label) {
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  if (row >= rows || !accepted[row])
    return;
  const uint64_t seg = row / nv;
  const int32_t u = static_cast<int32_t>(row - seg * nv);
  const uint32_t *acc = accepted + seg * nv;
  int32_t best = u;
  for (uint64_t j = nbr_off[u]; j < nbr_off[u + 1]; ++j) {
    const int32_t v = nbr_idx[j];
    if (v < best && v >= 0 && acc[v])
      best = v;
  }
  label[row] = best;
}

// k_cc_jump: one round of pointer jumping, parent <- grandparent (still an
// ancestor).  A few rounds after k_cc_init shorten the chains the
// initial parents form (along a grid they run a whole column) before the
// hooks walk them; the number of rounds does not change the result.
__global__ void __launch_bounds__(kVThreads)
k_cc_jump(const uint32_t *__restrict__ accepted, uint32_t nv, uint64_t rows, int32_t *label) {
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  if (row >= rows || !accepted[row])
    return;
  const uint64_t seg = row / nv;
  int32_t *par = label + seg * nv;
  const int32_t u = static_cast<int32_t>(row - seg * nv);
  const int32_t p = load_parent(par + u);
  const int32_t g = load_parent(par + p);
  if (g != p)
    store_parent(par + u, g);
}

// k_cc_hook: every ring edge (u, v), u < v, with both ends accepted joins
// their trees by hooking the larger root under the smaller with a
// compare-and-swap; a failed swap moves on to the root's new parent.
__global__ void __launch_bounds__(kVThreads)
k_cc_hook(const uint64_t *__restrict__ nbr_off, const int32_t *__restrict__ nbr_idx,
          const uint32_t *__restrict__ accepted, uint32_t nv, uint64_t rows, int32_t *label) {
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  if (row >= rows || !accepted[row])
    return;
  const uint64_t seg = row / nv;
  const int32_t u = static_cast<int32_t>(row - seg * nv);
  int32_t *par = label + seg * nv;
  const uint32_t *acc = accepted + seg * nv;
  for (uint64_t j = nbr_off[u]; j < nbr_off[u + 1]; ++j) {
    const int32_t v = nbr_idx[j];
    if (v <= u || static_cast<uint32_t>(v) >= nv || !acc[v])
      continue;
    int32_t a = cc_find(par, u), b = cc_find(par, v);
    while (a != b) {
      const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
      int32_t expect = hi;
      if (__hip_atomic_compare_exchange_strong(par + hi, &expect, lo, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        break;
      // hi was no longer a root: continue from its parent (smaller than hi)
      if (a == hi)
        a = expect;
      else
        b = expect;
    }
  }
}

// label[r] = root of r's vertex (the smallest member of its patch).
__global__ void __launch_bounds__(kVThreads)
k_cc_flatten(uint32_t nv, uint64_t rows, int32_t *label) {
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  if (row >= rows)
    return;
  const uint64_t seg = row / nv;
  const int32_t u = static_cast<int32_t>(row - seg * nv);
  int32_t *par = label + seg * nv;
  if (load_parent(par + u) < 0)
    return;
  store_parent(par + u, cc_find(par, u));
}

// Accepted rows in row order: vertex, patch root and squared distance.
__global__ void __launch_bounds__(kVThreads)
k_compact(const uint32_t *__restrict__ accepted, const uint64_t *__restrict__ pos,
          const int32_t *__restrict__ label, const double *__restrict__ dist2, uint32_t nv,
          uint64_t rows, int32_t *__restrict__ vid, int32_t *__restrict__ root,
          double *__restrict__ d2) {
  const uint64_t row = static_cast<uint64_t>(blockIdx.x) * kVThreads + threadIdx.x;
  if (row >= rows || !accepted[row])
    return;
  const uint64_t o = pos[row];
  const uint64_t seg = row / nv;
  vid[o] = static_cast<int32_t>(row - seg * nv);
  root[o] = label[row];
  d2[o] = dist2[row];
}

// Spreads the low 16 bits of v to every third bit (bit k -> bit 3k).
__device__ inline uint64_t spread3(uint32_t v) {
  uint64_t x = v & 0xffffu;
  x = (x | (x << 16)) & 0x0000ff0000ffull;
  x = (x | (x << 8)) & 0x00f00f00f00full;
  x = (x | (x << 4)) & 0x0c30c30c30c3ull;
  x = (x | (x << 2)) & 0x249249249249ull;
  return x;
}

// Morton key (16 bits per axis over the surface's box, x in the lowest bit)
// and vertex id of every vertex: sorted, the ids give the order in which a
// workgroup's threads take vertices, so the 64 rows of a wave lie close
// together and search similar parts of the tree.
__global__ void __launch_bounds__(kVThreads)
k_morton(const double *__restrict__ x, const double *__restrict__ y,
         const double *__restrict__ z, uint32_t nv, double lox, double loy, double loz,
         double scale, uint64_t *__restrict__ keys, uint32_t *__restrict__ ids) {
  const uint32_t i = blockIdx.x * kVThreads + threadIdx.x;
  if (i >= nv)
    return;
  auto q = [scale](double v, double lo) {
    const double t = (v - lo) * scale;
    return static_cast<uint32_t>(t <= 0.0 ? 0.0 : (t >= 65535.0 ? 65535.0 : t));
  };
  keys[i] = spread3(q(x[i], lox)) | (spread3(q(y[i], loy)) << 1) | (spread3(q(z[i], loz)) << 2);
  ids[i] = i;
}

// First compact position of every segment: bnd[k] = pos[k * nv], k <= segs.
__global__ void k_bounds(const uint64_t *__restrict__ pos, uint32_t nv, uint32_t segs,
                         uint64_t *__restrict__ bnd) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k <= segs)
    bnd[k] = pos[static_cast<uint64_t>(k) * nv];
}

}  // namespace

// ------------------------------------------------------------- host side
namespace {

struct Clusters {
  uint64_t nc;
  const uint8_t *reflection;
  const double *rotation9, *translation3, *plane_normal3, *plane_offset;
};

// Groups one segment's compact rows (ascending vertices with their patch
// roots) into patches: a root is its patch's smallest vertex, so patches
// appear in the order of their roots, which is the reference's order of
// breadth-first seeds (verification.cpp:27-47); members stay ascending.
// Writes the n members to `members` and appends the patch sizes to `sizes`;
// slot_of is scratch indexed by vertex (read only where written first).
void group_segment(const int32_t *vid, const int32_t *root, uint64_t n, int32_t *slot_of,
                   std::vector<uint64_t> &sizes, int32_t *members) {
  sizes.clear();
  for (uint64_t k = 0; k < n; ++k) {
    if (root[k] == vid[k]) {
      slot_of[vid[k]] = static_cast<int32_t>(sizes.size());
      sizes.push_back(0);
    }
    ++sizes[static_cast<size_t>(slot_of[root[k]])];
  }
  std::vector<uint64_t> next(sizes.size());
  uint64_t run = 0;
  for (size_t g = 0; g < sizes.size(); ++g) {
    next[g] = run;
    run += sizes[g];
  }
  for (uint64_t k = 0; k < n; ++k)
    members[next[static_cast<size_t>(slot_of[root[k]])]++] = vid[k];
}

// The surface (tree and leaf-ordered faces) of a mesh, kept on the session.
void surface_build(Session *s, const mvs_points *vertices, uint64_t nf, const int32_t *const f[3]) {
  const uint64_t nv = vertices->count;
  if (nv > static_cast<uint64_t>(INT32_MAX) || nf > static_cast<uint64_t>(INT32_MAX))
    raise(MVS_EOVERFLOW, "verification: vertex and face ids must fit int32");
  if (nf && (!vertices->x || !vertices->y || !vertices->z))
    raise(MVS_EINVAL, "verification: null vertex array");
  for (uint64_t i = 0; i < nf; ++i)
    for (int c = 0; c < 3; ++c)
      if (f[c][i] < 0 || static_cast<uint64_t>(f[c][i]) >= nv)
        raise(MVS_EINVAL, "verification: face vertex index out of range");
  s->ver_surface = false;
  TreeBuilder tb;
  tb.x = vertices->x, tb.y = vertices->y, tb.z = vertices->z;
  for (int c = 0; c < 3; ++c)
    tb.f[c] = f[c];
  for (int k = 0; k < 3; ++k)
    tb.cen[k].resize(nf);
  tb.fbox.resize(6 * nf);
  parallel_chunks(nf, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const int v0 = f[0][i], v1 = f[1][i], v2 = f[2][i];
      for (int k = 0; k < 3; ++k) {
        const double a = tb.coord(v0, k), bb = tb.coord(v1, k), c = tb.coord(v2, k);
        // the reference's centroid, (a + b + c) / 3 (surface.h:36-40)
        tb.cen[k][i] = (a + bb + c) / 3;
        tb.fbox[6 * i + k] = std::min(std::min(a, bb), c);
        tb.fbox[6 * i + 3 + k] = std::max(std::max(a, bb), c);
      }
    }
  });
  tb.ids.resize(nf);
  std::iota(tb.ids.begin(), tb.ids.end(), 0);
  const auto shape = tb.shape_of(static_cast<int>(nf));
  tb.nodes.resize(static_cast<size_t>(shape.first));
  tb.build(0, 0, static_cast<int>(nf), 4);
  // leaf-ordered faces (corners and the degenerate test of
  // src/core/math.cpp:74) and their unit normals (verification.cpp:71-73)
  std::vector<double> fr(std::max<uint64_t>(nf, 1) * kFaceWords, 0.0);
  std::vector<double> fn(std::max<uint64_t>(nf, 1) * 3, 0.0);
  parallel_chunks(nf, [&](uint64_t b, uint64_t e) {
    for (uint64_t k = b; k < e; ++k) {
      const int face = tb.ids[k];
      V3 c[3];
      for (int j = 0; j < 3; ++j) {
        const int v = f[j][face];
        c[j] = v3(tb.x[v], tb.y[v], tb.z[v]);
      }
      const V3 ab = c[1] - c[0], ac = c[2] - c[0];
      const V3 n = unit(cross(ab, ac));
      double *o = fr.data() + k * kFaceWords;
      for (int j = 0; j < 3; ++j)
        o[3 * j] = c[j].x, o[3 * j + 1] = c[j].y, o[3 * j + 2] = c[j].z;
      o[9] = dot(cross(ab, ac), cross(ab, ac)) < 1e-14 * 1e-14 ? 1.0 : 0.0;
      fn[3 * k] = n.x, fn[3 * k + 1] = n.y, fn[3 * k + 2] = n.z;
    }
  });
  // search records: inner nodes in pre-order, each with its children's boxes
  if (nf >= kMaxLeafStart)
    raise(MVS_EOVERFLOW, "verification: at most 2^27 faces");
  std::vector<int32_t> inner_of(tb.nodes.size(), -1);
  int32_t ninner = 0;
  for (size_t i = 0; i < tb.nodes.size(); ++i)
    if (tb.nodes[i].left >= 0)
      inner_of[i] = ninner++;
  auto ref_of = [&](int node) {
    const SNode &c = tb.nodes[static_cast<size_t>(node)];
    return c.left >= 0 ? inner_of[static_cast<size_t>(node)] : -1 - (16 * c.begin + (c.end - c.begin));
  };
  std::vector<Inner> inner(std::max<int32_t>(ninner, 1));
  for (size_t i = 0; i < tb.nodes.size(); ++i) {
    if (inner_of[i] < 0)
      continue;
    Inner &r = inner[static_cast<size_t>(inner_of[i])];
    std::memset(&r, 0, sizeof(Inner));
    const int kids[2] = {tb.nodes[i].left, tb.nodes[i].right};
    for (int c = 0; c < 2; ++c) {
      const SNode &k = tb.nodes[static_cast<size_t>(kids[c])];
      for (int a = 0; a < 3; ++a)
        r.lo[c][a] = k.lo[a], r.hi[c][a] = k.hi[a];
      r.ref[c] = ref_of(kids[c]);
    }
  }
  s->ver_root = ref_of(0);
  for (int k = 0; k < 3; ++k) {
    s->ver_box[k] = nf ? tb.nodes[0].lo[k] : 0.0;
    s->ver_box[3 + k] = nf ? tb.nodes[0].hi[k] : 0.0;
  }
  s->ver_stack = shape.second + 1;
  if (static_cast<size_t>(s->ver_stack) * kVThreads * (sizeof(int) + sizeof(float)) > 64 * 1024)
    raise(MVS_EOVERFLOW, "verification: surface tree too deep");
  // queued work may still read the previous surface
  sync(s);
  s->ver_nodes.alloc(inner.size() * sizeof(Inner));
  s->ver_faces.alloc(fr.size() * sizeof(double));
  s->ver_normals.alloc(fn.size() * sizeof(double));
  copy_h2d(s, s->ver_nodes.ptr, inner.data(), inner.size() * sizeof(Inner));
  copy_h2d(s, s->ver_faces.ptr, fr.data(), fr.size() * sizeof(double));
  copy_h2d(s, s->ver_normals.ptr, fn.data(), fn.size() * sizeof(double));
  sync(s);  // the host vectors go out of scope
  s->ver_surface_nv = nv;
  s->ver_surface = true;
}

// Rows, patches and compaction on the device; sides, patches, rms and
// coverage assembled on the host in the reference's order
// (verification.cpp:57-107).  pts: 6 nv device doubles (x, y, z, normal
// x, y, z); neighbours on the device; weights on the host.
void run(Session *s, uint64_t nv, const double *pts, const uint64_t *noff, const int32_t *nidx,
         const double *weights, const Clusters &cl, double verify_threshold, double box,
         mvs_verification_counts *counts) {
  if (!s->ver_surface)
    raise(MVS_EINVAL, "verification: no surface (mvs_verification_surface)");
  if (nv != s->ver_surface_nv)
    raise(MVS_EINVAL, "verification: vertex count differs from the surface's");
  if (nv > static_cast<uint64_t>(INT32_MAX))
    raise(MVS_EOVERFLOW, "verification: vertex ids must fit int32");
  const uint64_t nc = cl.nc;
  if (nc > UINT32_MAX / 4)
    raise(MVS_EOVERFLOW, "verification: too many clusters");
  if (nv && !weights)
    raise(MVS_EINVAL, "verification: null weights");
  if (nc && (!cl.reflection || !cl.rotation9 || !cl.translation3 || !cl.plane_normal3 ||
             !cl.plane_offset))
    raise(MVS_EINVAL, "verification: null cluster array");
  const double threshold = verify_threshold * box;
  const uint64_t segs = 2 * nc, rows = segs * nv;
  if (nv && segs > (static_cast<uint64_t>(UINT32_MAX) * kVThreads) / nv)
    raise(MVS_EOVERFLOW, "verification: clusters x vertices exceed one launch");
  for (int sd = 0; sd < 2; ++sd) {
    s->ver_side_off[sd].assign(1, 0);
    s->ver_side[sd].clear();
    s->ver_groups[sd].assign(1, 0);
    s->ver_patch_off[sd].assign(1, 0);
    s->ver_members[sd].clear();
  }
  s->ver_cov.assign(nc, 0.0);
  s->ver_rms.assign(nc, 0.0);
  s->ver_rows = 0;

  std::vector<uint64_t> bnd(segs + 1, 0);
  const int32_t *h_vid = nullptr, *h_root = nullptr;
  const double *h_d2 = nullptr;
  const auto t_start = std::chrono::steady_clock::now();
  auto t_dev = t_start;
  if (rows > 0) {
    std::vector<Motion> mo(nc);
    for (uint64_t c = 0; c < nc; ++c) {
      std::memset(&mo[c], 0, sizeof(Motion));
      for (int k = 0; k < 9; ++k)
        mo[c].r[k] = cl.rotation9[9 * c + k];
      for (int k = 0; k < 3; ++k) {
        mo[c].t[k] = cl.translation3[3 * c + k];
        mo[c].n[k] = cl.plane_normal3[3 * c + k];
      }
      mo[c].off = cl.plane_offset[c];
      mo[c].reflection = cl.reflection[c] ? 1 : 0;
    }
    const size_t lds = static_cast<size_t>(s->ver_stack) * kVThreads * (sizeof(int) + sizeof(float));
    Motion *d_mo = static_cast<Motion *>(scratch(s, kSlotVerMotions, nc * sizeof(Motion)));
    uint32_t *d_acc = scratch_as<uint32_t>(s, kSlotVerAccepted, rows);
    int32_t *d_label = scratch_as<int32_t>(s, kSlotVerLabel, rows);
    double *d_dist = scratch_as<double>(s, kSlotVerDist, rows);
    uint64_t *d_pos = scratch_as<uint64_t>(s, kSlotVerPos, rows + 1);
    uint64_t *d_bnd = scratch_as<uint64_t>(s, kSlotVerBounds, segs + 1);
    auto *d_stats = scratch_as<unsigned long long>(s, kSlotVerStats, 3);
    copy_h2d(s, d_mo, mo.data(), nc * sizeof(Motion));  // mo outlives the sync below
    zero_device(s, d_stats, 3 * sizeof(unsigned long long));

    const unsigned grid = blocks_for(rows, kVThreads);
    const uint32_t nv32 = static_cast<uint32_t>(nv);
    // rows taken in the Morton order of their vertices
    uint64_t *d_keys = scratch_as<uint64_t>(s, kSlotVerKeys, nv);
    uint64_t *d_keys2 = scratch_as<uint64_t>(s, kSlotVerKeys2, nv);
    uint32_t *d_ids = scratch_as<uint32_t>(s, kSlotVerIds, nv);
    uint32_t *d_order = scratch_as<uint32_t>(s, kSlotVerOrder, nv);
    {
      const double *lo = s->ver_box, *hi = s->ver_box + 3;
      const double ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
      const double scale = ext > 0 ? 65535.0 / ext : 0.0;
      MVS_LAUNCH(s, "verify_morton", k_morton, blocks_for(nv, kVThreads), kVThreads, 0, pts,
                 pts + nv, pts + 2 * nv, nv32, lo[0], lo[1], lo[2], scale, d_keys, d_ids);
      radix_sort_pairs_u64(s, d_keys, d_ids, d_keys2, d_order, nv, 48);
    }
    MVS_LAUNCH(s, "verify_rows", k_verify, grid, kVThreads, lds, pts, pts + nv, pts + 2 * nv,
               pts + 3 * nv, pts + 4 * nv, pts + 5 * nv, nv32, rows,
               reinterpret_cast<const Inner *>(s->ver_nodes.ptr), s->ver_root,
               reinterpret_cast<const double *>(s->ver_faces.ptr),
               reinterpret_cast<const double *>(s->ver_normals.ptr), d_mo, threshold,
               s->ver_stack, d_order, d_acc,
               d_label, d_dist, d_stats);
    MVS_LAUNCH(s, "verify_cc_init", k_cc_init, grid, kVThreads, 0, noff, nidx, d_acc, nv32, rows,
               d_label);
    for (int round = 0; round < kJumpRounds; ++round)
      MVS_LAUNCH(s, "verify_cc_jump", k_cc_jump, grid, kVThreads, 0, d_acc, nv32, rows, d_label);
    MVS_LAUNCH(s, "verify_cc_hook", k_cc_hook, grid, kVThreads, 0, noff, nidx, d_acc, nv32, rows,
               d_label);
    MVS_LAUNCH(s, "verify_cc_flatten", k_cc_flatten, grid, kVThreads, 0, nv32, rows, d_label);
    exclusive_scan_u32(s, d_acc, d_pos, rows);
    MVS_LAUNCH(s, "verify_bounds", k_bounds, blocks_for(segs + 1, 64), 64, 0, d_pos, nv32,
               static_cast<uint32_t>(segs), d_bnd);
    copy_d2h(s, bnd.data(), d_bnd, (segs + 1) * sizeof(uint64_t));
    const uint64_t total = bnd[segs];
    int32_t *d_vid = scratch_as<int32_t>(s, kSlotVerVid, std::max<uint64_t>(total, 1));
    int32_t *d_root = scratch_as<int32_t>(s, kSlotVerRoot, std::max<uint64_t>(total, 1));
    double *d_d2 = scratch_as<double>(s, kSlotVerD2, std::max<uint64_t>(total, 1));
    MVS_LAUNCH(s, "verify_compact", k_compact, grid, kVThreads, 0, d_acc, d_pos, d_label,
               d_dist, nv32, rows, d_vid, d_root, d_d2);
    // the compact rows through pinned staging: ids, roots, squared distances
    const uint64_t tb = std::max<uint64_t>(total, 1);
    auto *st = static_cast<unsigned char *>(host_stage(s, 16 * tb));
    int32_t *sv = reinterpret_cast<int32_t *>(st), *sr = sv + tb;
    double *sd2 = reinterpret_cast<double *>(st + 8 * tb);
    t_dev = std::chrono::steady_clock::now();
    copy_d2h(s, sv, d_vid, total * sizeof(int32_t));
    copy_d2h(s, sr, d_root, total * sizeof(int32_t));
    copy_d2h(s, sd2, d_d2, total * sizeof(double));
    h_vid = sv, h_root = sr, h_d2 = sd2;
    s->ver_rows = rows;
  }
  const auto t_copy = std::chrono::steady_clock::now();

  double total_area = 0;
  for (uint64_t i = 0; i < nv; ++i)
    total_area += weights[i];
  for (int sd = 0; sd < 2; ++sd) {
    auto &off = s->ver_side_off[sd];
    off.assign(nc + 1, 0);
    for (uint64_t c = 0; c < nc; ++c)
      off[c + 1] = off[c] + (bnd[2 * c + sd + 1] - bnd[2 * c + sd]);
    s->ver_side[sd].resize(off[nc]);
    s->ver_members[sd].resize(off[nc]);
  }
  // clusters in parallel; within a cluster the reference's order: the error
  // over side A, the area over side A then side B (verification.cpp:57-96)
  std::vector<std::vector<uint64_t>> sizes(2 * nc);
  const unsigned workers = static_cast<unsigned>(std::max<uint64_t>(
      1, std::min<uint64_t>(nc, std::min(16u, std::max(1u, std::thread::hardware_concurrency())))));
  if (s->ver_slot.size() < workers)
    s->ver_slot.resize(workers);
  std::atomic<uint64_t> next{0};
  auto work = [&](unsigned w) {
    auto &slot = s->ver_slot[w];
    if (slot.size() < nv)
      slot.resize(nv);
    for (uint64_t c; (c = next.fetch_add(1)) < nc;) {
      double error = 0, area = 0;
      for (int sd = 0; sd < 2; ++sd) {
        const uint64_t b = bnd[2 * c + sd], e = bnd[2 * c + sd + 1];
        const uint64_t o = s->ver_side_off[sd][c];
        if (e > b)
          std::memcpy(s->ver_side[sd].data() + o, h_vid + b, (e - b) * sizeof(int32_t));
        for (uint64_t k = b; k < e; ++k) {
          if (sd == 0)
            error += h_d2[k];
          area += weights[h_vid[k]];
        }
        group_segment(h_vid + b, h_root + b, e - b, slot.data(), sizes[2 * c + sd],
                      s->ver_members[sd].data() + o);
      }
      const uint64_t na = bnd[2 * c + 1] - bnd[2 * c];
      s->ver_rms[c] = na == 0 ? 0 : std::sqrt(error / static_cast<double>(na));
      s->ver_cov[c] = total_area > 0 ? std::min(1.0, area / total_area) : 0;
    }
  };
  {
    std::vector<std::thread> ts;
    for (unsigned w = 1; w < workers; ++w)
      ts.emplace_back(work, w);
    work(0);
    for (auto &t : ts)
      t.join();
  }
  for (int sd = 0; sd < 2; ++sd) {
    auto &po = s->ver_patch_off[sd];
    auto &gr = s->ver_groups[sd];
    for (uint64_t c = 0; c < nc; ++c) {
      for (uint64_t z : sizes[2 * c + sd])
        po.push_back(po.back() + z);
      gr.push_back(po.size() - 1);
    }
  }
  const auto t_host = std::chrono::steady_clock::now();
  auto us = [](auto a, auto b) {
    return static_cast<uint64_t>(std::chrono::duration<double, std::micro>(b - a).count());
  };
  s->ver_us[0] = us(t_start, t_dev);
  s->ver_us[1] = us(t_dev, t_copy);
  s->ver_us[2] = us(t_copy, t_host);
  counts->side_a = s->ver_side[0].size();
  counts->side_b = s->ver_side[1].size();
  counts->patches_a = s->ver_patch_off[0].size() - 1;
  counts->patches_b = s->ver_patch_off[1].size() - 1;
}

void put_side(Session *s, int sd, const mvs_verification_side *o) {
  if (!o)
    return;
  auto put = [](auto *dst, const auto &v) {
    if (dst && !v.empty())
      std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  put(o->side_offsets, s->ver_side_off[sd]);
  put(o->side, s->ver_side[sd]);
  put(o->patch_groups, s->ver_groups[sd]);
  put(o->patch_offsets, s->ver_patch_off[sd]);
  put(o->patch_members, s->ver_members[sd]);
}

}  // namespace
}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_verification_surface(mvs_session *sp, const mvs_points *vertices, uint64_t face_count,
                             const int32_t *f0, const int32_t *f1, const int32_t *f2) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!vertices)
      raise(MVS_EINVAL, "verification: null vertices");
    if (face_count && (!f0 || !f1 || !f2))
      raise(MVS_EINVAL, "verification: null face array");
    const int32_t *f[3] = {f0, f1, f2};
    surface_build(s, vertices, face_count, f);
  });
}

int mvs_dev_verification(mvs_session *sp, uint64_t nv, const double *pts6,
                         const uint64_t *neighbor_offsets, const int32_t *neighbor_indices,
                         const double *weights, uint64_t cluster_count, const uint8_t *reflection,
                         const double *rotation9, const double *translation3,
                         const double *plane_normal3, const double *plane_offset,
                         double verify_threshold, double box, mvs_verification_counts *counts) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!counts || (nv && (!pts6 || !neighbor_offsets || !neighbor_indices)))
      raise(MVS_EINVAL, "verification: null argument");
    run(s, nv, pts6, neighbor_offsets, neighbor_indices, weights,
        Clusters{cluster_count, reflection, rotation9, translation3, plane_normal3, plane_offset},
        verify_threshold, box, counts);
  });
}

int mvs_verification(mvs_session *sp, const mvs_points *vertices, uint64_t face_count,
                     const int32_t *f0, const int32_t *f1, const int32_t *f2,
                     const mvs_points *normals, const double *weights,
                     const uint64_t *neighbor_offsets, const int32_t *neighbor_indices,
                     uint64_t cluster_count, const uint8_t *reflection, const double *rotation9,
                     const double *translation3, const double *plane_normal3,
                     const double *plane_offset, double verify_threshold, double box,
                     mvs_verification_counts *counts) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!vertices || !normals || !counts || !neighbor_offsets)
      raise(MVS_EINVAL, "verification: null argument");
    const uint64_t nv = vertices->count;
    if (normals->count != nv)
      raise(MVS_EINVAL, "verification: normals and vertices differ in size");
    if (nv && (!vertices->x || !vertices->y || !vertices->z || !normals->x || !normals->y ||
               !normals->z || !weights))
      raise(MVS_EINVAL, "verification: null vertex array");
    if (face_count && (!f0 || !f1 || !f2))
      raise(MVS_EINVAL, "verification: null face array");
    if (neighbor_offsets[0] != 0)
      raise(MVS_EINVAL, "verification: neighbor offsets must start at 0");
    for (uint64_t i = 0; i < nv; ++i)
      if (neighbor_offsets[i + 1] < neighbor_offsets[i])
        raise(MVS_EINVAL, "verification: neighbor offsets must not decrease");
    const uint64_t nnz = nv ? neighbor_offsets[nv] : 0;
    if (nnz && !neighbor_indices)
      raise(MVS_EINVAL, "verification: null neighbor indices");
    for (uint64_t j = 0; j < nnz; ++j)
      if (neighbor_indices[j] < 0 || static_cast<uint64_t>(neighbor_indices[j]) >= nv)
        raise(MVS_EINVAL, "verification: neighbor index out of range");
    const int32_t *f[3] = {f0, f1, f2};
    surface_build(s, vertices, face_count, f);
    double *d_pts = scratch_as<double>(s, kSlotVerPoints, std::max<uint64_t>(6 * nv, 1));
    uint64_t *d_noff = scratch_as<uint64_t>(s, kSlotVerNbrOff, nv + 1);
    int32_t *d_nidx = scratch_as<int32_t>(s, kSlotVerNbrIdx, std::max<uint64_t>(nnz, 1));
    if (cluster_count) {
      const double *src[6] = {vertices->x, vertices->y, vertices->z,
                              normals->x,  normals->y,  normals->z};
      for (int k = 0; k < 6; ++k)
        copy_h2d(s, d_pts + k * nv, src[k], nv * sizeof(double));
      copy_h2d(s, d_noff, neighbor_offsets, (nv + 1) * sizeof(uint64_t));
      copy_h2d(s, d_nidx, neighbor_indices, nnz * sizeof(int32_t));
    }
    run(s, nv, d_pts, d_noff, d_nidx, weights,
        Clusters{cluster_count, reflection, rotation9, translation3, plane_normal3, plane_offset},
        verify_threshold, box, counts);
  });
}

int mvs_verification_fetch(mvs_session *sp, const mvs_verification_side *side_a,
                           const mvs_verification_side *side_b, double *coverage, double *rms) {
  return guarded([&] {
    Session *s = as_session(sp);
    put_side(s, 0, side_a);
    put_side(s, 1, side_b);
    if (coverage && !s->ver_cov.empty())
      std::memcpy(coverage, s->ver_cov.data(), s->ver_cov.size() * sizeof(double));
    if (rms && !s->ver_rms.empty())
      std::memcpy(rms, s->ver_rms.data(), s->ver_rms.size() * sizeof(double));
  });
}

int mvs_verification_stats(mvs_session *sp, uint64_t out[6]) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!out)
      raise(MVS_EINVAL, "verification stats: null output");
    for (int k = 0; k < 6; ++k)
      out[k] = k < 3 ? 0 : s->ver_us[k - 3];
    if (!s->ver_rows)
      return;
    copy_d2h(s, out, scratch(s, kSlotVerStats, 3 * sizeof(uint64_t)), 3 * sizeof(uint64_t));
  });
}

int mvs_verification_rows(mvs_session *sp, uint8_t *accepted, double *dist2) {
  return guarded([&] {
    Session *s = as_session(sp);
    const uint64_t rows = s->ver_rows;
    if (!rows)
      return;
    if (accepted) {
      std::vector<uint32_t> a(rows);
      copy_d2h(s, a.data(), scratch(s, kSlotVerAccepted, rows * sizeof(uint32_t)),
               rows * sizeof(uint32_t));
      for (uint64_t r = 0; r < rows; ++r)
        accepted[r] = static_cast<uint8_t>(a[r]);
    }
    if (dist2)
      copy_d2h(s, dist2, scratch(s, kSlotVerDist, rows * sizeof(double)),
               rows * sizeof(double));
  });
}

}  // extern "C"
