// Mode refinement: closest points on the mesh through a BVH, and the
// point-to-plane alignment energy of a transform with its adjoint (gradient)
// and Gauss-Newton matrix.  No reference counterpart for the energy; the
// closest-point query follows the reference's surface queries
// (src/cpu/surface.h:12-96, src/core/math.cpp:72-112): a median-split BVH
// with at most 8 faces per leaf and the exact point-triangle distance.
//
// Coordinates are relative to the model centre (FP32).
//   rigid:   T(p) = R(w) (R0 p + t0) + v,  parameters (w, v), R(w) = exp([w]x)
//   mirror:  T(p) = p - 2 (u.p + d) u,     u = m / |m|, parameters (m, d)
// Residual of point j: r_j = n_j . (T(p_j) - q_j) with q_j, n_j the closest
// mesh point of the current transform and its face normal (held fixed
// while the parameters move).  E = sum r^2, g = dE/dtheta,
// H = 2 sum J_j^T J_j (J_j = dr_j/dtheta).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "pairing.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr int kLeaf = 8;
constexpr uint64_t kMaxFaces = 1ULL << 27;  // leaf code first * 16 + count < 2^31

// Inner node of the tree: the boxes and references of both children in one
// 64-byte line, so a visit is one load.  A reference r >= 0 names an inner
// record; r < 0 a leaf, with code -1 - r = first * 16 + count (count <= 8
// faces from `first` in leaf order).
struct alignas(16) Inner {
  float lo0[3];
  int ref0;
  float hi0[3];
  int ref1;
  float lo1[3];
  int pad0;
  float hi1[3];
  int pad1;
};
static_assert(sizeof(Inner) == 64, "one 64-byte line per inner node");

// Face in leaf order: corners a, b, c and the unit normal (48 bytes).
struct alignas(16) Face {
  float a[3], b[3], c[3], n[3];
};
static_assert(sizeof(Face) == 48, "three 16-byte loads per face");

__device__ inline float box_dist2(const float *lo, const float *hi, float px, float py,
                                  float pz) {
  float dx = fmaxf(fmaxf(lo[0] - px, 0.f), px - hi[0]);
  float dy = fmaxf(fmaxf(lo[1] - py, 0.f), py - hi[1]);
  float dz = fmaxf(fmaxf(lo[2] - pz, 0.f), pz - hi[2]);
  return dx * dx + dy * dy + dz * dz;
}

struct P3 {
  float x, y, z;
};
__device__ inline P3 p3(float x, float y, float z) { return {x, y, z}; }
__device__ inline P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline P3 mul(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline float dot3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline P3 cross3(P3 a, P3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Closest point of segment ab to p.
__device__ inline P3 closest_on_segment(P3 p, P3 a, P3 b) {
  P3 d = sub(b, a);
  float l = dot3(d, d);
  float t = l > 0.f ? fminf(fmaxf(dot3(sub(p, a), d) / l, 0.f), 1.f) : 0.f;
  return add(a, mul(d, t));
}

// Closest point of triangle abc to p by Voronoi regions (Ericson, Real-Time
// Collision Detection, 2005, section 5.1.5); a degenerate triangle falls back
// to its edges.
__device__ inline P3 closest_on_triangle(P3 p, P3 a, P3 b, P3 c) {
  P3 ab = sub(b, a), ac = sub(c, a);
  P3 nn = cross3(ab, ac);
  if (dot3(nn, nn) < 1e-24f) {
    P3 q0 = closest_on_segment(p, a, b), q1 = closest_on_segment(p, a, c),
       q2 = closest_on_segment(p, b, c);
    float e0 = dot3(sub(p, q0), sub(p, q0)), e1 = dot3(sub(p, q1), sub(p, q1)),
          e2 = dot3(sub(p, q2), sub(p, q2));
    return e0 <= e1 && e0 <= e2 ? q0 : (e1 <= e2 ? q1 : q2);
  }
  P3 ap = sub(p, a);
  float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f)
    return a;
  P3 bp = sub(p, b);
  float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.f && d4 <= d3)
    return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f)
    return add(a, mul(ab, d1 / (d1 - d3)));
  P3 cp = sub(p, c);
  float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.f && d5 <= d6)
    return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f)
    return add(a, mul(ac, d2 / (d2 - d6)));
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f)
    return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  float den = 1.f / (va + vb + vc);
  return add(a, add(mul(ab, vb * den), mul(ac, vc * den)));
}

__device__ inline float face_d2(P3 p, const Face &f, P3 &c) {
  c = closest_on_triangle(p, p3(f.a[0], f.a[1], f.a[2]), p3(f.b[0], f.b[1], f.b[2]),
                          p3(f.c[0], f.c[1], f.c[2]));
  const P3 e = sub(p, c);
  return dot3(e, e);
}

// Nearest surface point to p: depth-first, nearer child first (the left one
// on a tie), subtrees whose box is farther than the best distance skipped.
// A stack entry carries its box distance, so an entry that the best hit has
// overtaken is dropped without a load.  The search alternates between
// descending to the next leaf and scanning it, so the lanes of a wave do the
// same kind of work at a time.  stk, bnd: this thread's columns of the LDS
// stack (stride kThreads, depth + 1 entries: one waiting sibling per level
// plus the two children just pushed).
struct Hit {
  P3 q;
  float d2;
  int k;  // leaf-ordered face, -1 when none
};
// h: the best hit known before the search (d2 = INFINITY: none).
__device__ inline Hit nearest(const Inner *inner, const Face *faces, int root, P3 p, int *stk,
                              float *bnd, Hit h) {
  int top = 1;
  stk[0] = root;
  bnd[0] = -INFINITY;
  while (true) {
    int leaf = 0;
    while (top > 0) {
      --top;
      const int ref = stk[top * kThreads];
      if (bnd[top * kThreads] > h.d2)
        continue;
      if (ref < 0) {
        leaf = ref;
        break;
      }
      const Inner nd = inner[ref];
      const float d0 = box_dist2(nd.lo0, nd.hi0, p.x, p.y, p.z),
                  d1 = box_dist2(nd.lo1, nd.hi1, p.x, p.y, p.z);
      // push the farther child first so the nearer one is searched first
      const bool swap = d0 > d1;
      const int near = swap ? nd.ref1 : nd.ref0, far = swap ? nd.ref0 : nd.ref1;
      const float dnear = swap ? d1 : d0, dfar = swap ? d0 : d1;
      if (dfar <= h.d2) {
        stk[top * kThreads] = far;
        bnd[top * kThreads] = dfar;
        ++top;
      }
      if (dnear <= h.d2) {
        stk[top * kThreads] = near;
        bnd[top * kThreads] = dnear;
        ++top;
      }
    }
    if (leaf == 0)
      break;
    const int code = -1 - leaf, first = code >> 4, end = first + (code & 15);
    for (int k = first; k < end; ++k) {
      const Face f = faces[k];
      // the box of the corners bounds the face's distance from below: a face
      // whose box is farther than the best hit (by a 2^-10 relative margin,
      // far above the FP32 rounding of either distance) skips the exact test
      const float lo[3] = {fminf(fminf(f.a[0], f.b[0]), f.c[0]),
                           fminf(fminf(f.a[1], f.b[1]), f.c[1]),
                           fminf(fminf(f.a[2], f.b[2]), f.c[2])};
      const float hi[3] = {fmaxf(fmaxf(f.a[0], f.b[0]), f.c[0]),
                           fmaxf(fmaxf(f.a[1], f.b[1]), f.c[1]),
                           fmaxf(fmaxf(f.a[2], f.b[2]), f.c[2])};
      if (box_dist2(lo, hi, p.x, p.y, p.z) > h.d2 * (1.f + 0x1p-10f))
        continue;
      P3 c;
      const float d = face_d2(p, f, c);
      if (d < h.d2) {
        h.d2 = d;
        h.q = c;
        h.k = k;
      }
    }
  }
  return h;
}

// The LDS stack of a block: `depth` references, then `depth` bounds, per
// thread (dynamic shared memory of depth * kThreads * 8 bytes).
struct Stack {
  int *stk;
  float *bnd;
  __device__ explicit Stack(int depth) {
    extern __shared__ int stack_lds[];
    stk = stack_lds + threadIdx.x;
    bnd = reinterpret_cast<float *>(stack_lds + depth * kThreads) + threadIdx.x;
  }
};

// One thread per query point (3 x n SoA).
__global__ void __launch_bounds__(kThreads)
    k_closest(const Inner *inner, const Face *faces, int root, int depth,
              const int32_t *face_id, uint64_t n, const float *pts, float *q, float *nrm,
              float *dist2, int32_t *face) {
  const Stack st(depth);
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const P3 p = p3(pts[i], pts[n + i], pts[2 * n + i]);
  const Hit h = nearest(inner, faces, root, p, st.stk, st.bnd, Hit{p, INFINITY, -1});
  q[i] = h.q.x, q[n + i] = h.q.y, q[2 * n + i] = h.q.z;
  if (h.k >= 0) {
    const float *fn = faces[h.k].n;
    nrm[i] = fn[0], nrm[n + i] = fn[1], nrm[2 * n + i] = fn[2];
  } else {
    nrm[i] = nrm[n + i] = nrm[2 * n + i] = 0.f;
  }
  dist2[i] = h.d2;
  if (face)
    face[i] = h.k >= 0 ? face_id[h.k] : -1;
}

// Rodrigues coefficients of R(w) = I + A K + B K^2 (K = [w]x, t = |w|),
// A = sin t / t, B = (1 - cos t) / t^2, and their derivatives divided by t,
// A'/t and B'/t; Taylor series below t = 0.1, where the closed forms cancel
// in FP32.
__device__ inline void rodrigues(float t, float &A, float &B, float &dA, float &dB) {
  const float t2 = t * t, t4 = t2 * t2;
  if (t < 0.1f) {
    A = 1.f - t2 / 6.f + t4 / 120.f;
    B = 0.5f - t2 / 24.f + t4 / 720.f;
    dA = -1.f / 3.f + t2 / 30.f - t4 / 840.f;
    dB = -1.f / 12.f + t2 / 180.f - t4 / 6720.f;
  } else {
    const float s = sinf(t), c = cosf(t);
    A = s / t;
    B = (1.f - c) / t2;
    dA = (t * c - s) / (t2 * t);
    dB = (t * s - 2.f * (1.f - c)) / (t2 * t2);
  }
}

// Transform of mode m applied to p (the energy's T, in FP32).
template <int KIND> struct ModeTransform {
  float M[9], t[3];
  __device__ ModeTransform(const float *theta, const float *base, uint32_t m) {
    if (KIND == 0) {
      const float *th = theta + 6 * static_cast<uint64_t>(m);
      const float *b = base + 12 * static_cast<uint64_t>(m);
      const P3 w = p3(th[0], th[1], th[2]);
      float A, B, dA, dB;
      rodrigues(sqrtf(dot3(w, w)), A, B, dA, dB);
      // R(w) = I + A K + B K^2, K = [w]x; M = R(w) R0, t = R(w) t0 + v
      const float K[9] = {0.f, -w.z, w.y, w.z, 0.f, -w.x, -w.y, w.x, 0.f};
      float Rw[9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          float kk = 0.f;
          for (int e = 0; e < 3; ++e)
            kk += K[3 * r + e] * K[3 * e + c];
          Rw[3 * r + c] = (r == c ? 1.f : 0.f) + A * K[3 * r + c] + B * kk;
        }
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
          M[3 * r + c] = Rw[3 * r] * b[c] + Rw[3 * r + 1] * b[3 + c] + Rw[3 * r + 2] * b[6 + c];
        t[r] = Rw[3 * r] * b[9] + Rw[3 * r + 1] * b[10] + Rw[3 * r + 2] * b[11] + th[3 + r];
      }
    } else {
      const float *th = theta + 4 * static_cast<uint64_t>(m);
      const P3 mm = p3(th[0], th[1], th[2]);
      const P3 u = mul(mm, 1.f / sqrtf(dot3(mm, mm)));
      const float uu[3] = {u.x, u.y, u.z};
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
          M[3 * r + c] = (r == c ? 1.f : 0.f) - 2.f * uu[r] * uu[c];
        t[r] = -2.f * th[3] * uu[r];
      }
    }
  }
  __device__ P3 operator()(P3 p) const {
    return p3(M[0] * p.x + M[1] * p.y + M[2] * p.z + t[0],
              M[3] * p.x + M[4] * p.y + M[5] * p.z + t[1],
              M[6] * p.x + M[7] * p.y + M[8] * p.z + t[2]);
  }
};

// Closest points of every mode's transformed points in one launch: thread
// (m, j) maps point j of mode m (pts: modes x 3 x npts, or one shared
// 3 x npts set when pts_stride == 0) and writes q, nrm (modes x 3 x npts)
// and dist2 (modes x npts).  Blocks never span two modes.
template <int KIND>
__global__ void __launch_bounds__(kThreads)
    k_closest_modes(const Inner *inner, const Face *faces, int root, int depth, uint32_t modes,
                    uint32_t npts, unsigned blocks_per_mode, const float *theta,
                    const float *base, const float *pts, uint64_t pts_stride, float *q,
                    float *nrm, float *dist2, int32_t *face, const int32_t *face_id,
                    const int32_t *leaf_of, uint32_t nfaces) {
  const Stack st(depth);
  const uint32_t m = blockIdx.x / blocks_per_mode;
  const uint64_t j = static_cast<uint64_t>(blockIdx.x % blocks_per_mode) * kThreads + threadIdx.x;
  if (m >= modes || j >= npts)
    return;
  const ModeTransform<KIND> T(theta, base, m);
  const float *pm = pts + m * pts_stride;
  const P3 p = T(p3(pm[j], pm[npts + j], pm[2 * static_cast<uint64_t>(npts) + j]));
  const uint64_t fo = static_cast<uint64_t>(m) * npts + j;
  // warm start: the face given for this point (the previous step's) bounds
  // the search from the start; the tree then finds anything nearer
  Hit init{p, INFINITY, -1};
  if (face) {
    const int32_t f = face[fo];
    if (f >= 0 && static_cast<uint32_t>(f) < nfaces) {
      const int k = leaf_of[f];
      P3 c;
      const float d = face_d2(p, faces[k], c);
      init = Hit{c, d, k};
    }
  }
  const Hit h = nearest(inner, faces, root, p, st.stk, st.bnd, init);
  if (face)
    face[fo] = h.k >= 0 ? face_id[h.k] : -1;
  const uint64_t o = static_cast<uint64_t>(m) * 3 * npts + j;
  q[o] = h.q.x, q[o + npts] = h.q.y, q[o + 2 * static_cast<uint64_t>(npts)] = h.q.z;
  const float *f = h.k >= 0 ? faces[h.k].n : nullptr;
  nrm[o] = f ? f[0] : 0.f, nrm[o + npts] = f ? f[1] : 0.f;
  nrm[o + 2 * static_cast<uint64_t>(npts)] = f ? f[2] : 0.f;
  dist2[fo] = h.d2;
}

template <int NP> __device__ inline void block_sum(float *v, float *sh) {
  // v[0..NP) summed over the block into v (every thread), fixed order
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int k = 0; k < NP; ++k)
    for (int o = kWave / 2; o > 0; o >>= 1)
      v[k] += __shfl_xor(v[k], o, kWave);
  if (lane == 0)
    for (int k = 0; k < NP; ++k)
      sh[wave * NP + k] = v[k];
  __syncthreads();
  for (int k = 0; k < NP; ++k) {
    float s = 0.f;
    for (int w = 0; w < kThreads / kWave; ++w)
      s += sh[w * NP + k];
    v[k] = s;
  }
}

// Points of one mode per block of the energy kernel (four per thread).
constexpr uint32_t kAlignChunk = 4 * kThreads;

template <int KIND> struct AlignSizes {
  static constexpr int NP = KIND == 0 ? 6 : 4;  // parameters
  static constexpr int NH = NP * (NP + 1) / 2;  // upper triangle of H
  static constexpr int NV = 1 + NP + NH;        // E, g, H
};

// Block (m, c) of `chunks` per mode: the residuals of points
// [c kAlignChunk, (c + 1) kAlignChunk) of mode m, and their E, g (NP) and
// upper triangle of H (NP (NP + 1) / 2) summed in FP32 in a fixed order into
// part[(m chunks + c) NV ...]; k_align_sum adds the chunks in order.
template <int KIND>
__global__ void __launch_bounds__(kThreads)
    k_align(uint32_t modes, uint32_t npts, uint32_t chunks, const float *theta,
            const float *base, const float *pts, const float *q, const float *nrm,
            float *part) {
  constexpr int NP = AlignSizes<KIND>::NP, NV = AlignSizes<KIND>::NV;
  __shared__ float sh[(kThreads / kWave) * NV];
  const uint32_t m = blockIdx.x / chunks, c = blockIdx.x % chunks;
  if (m >= modes)
    return;
  const float *th = theta + static_cast<uint64_t>(m) * NP;
  const uint64_t off = static_cast<uint64_t>(m) * npts * 3;
  const float *px = pts + off, *py = px + npts, *pz = py + npts;
  const float *qx = q + off, *qy = qx + npts, *qz = qy + npts;
  const float *nx = nrm + off, *ny = nx + npts, *nz = ny + npts;
  float acc[NV];
  for (int k = 0; k < NV; ++k)
    acc[k] = 0.f;
  // per-mode constants
  P3 w{}, v{}, u{};
  float A = 1, B = 0.5f, dA = 0, dB = 0, R0[9] = {}, t0[3] = {}, d = 0, inv_m = 1;
  if (KIND == 0) {
    w = p3(th[0], th[1], th[2]);
    v = p3(th[3], th[4], th[5]);
    rodrigues(sqrtf(dot3(w, w)), A, B, dA, dB);
    for (int k = 0; k < 9; ++k)
      R0[k] = base[12 * m + k];
    for (int k = 0; k < 3; ++k)
      t0[k] = base[12 * m + 9 + k];
  } else {
    P3 mm = p3(th[0], th[1], th[2]);
    inv_m = 1.f / sqrtf(dot3(mm, mm));
    u = mul(mm, inv_m);
    d = th[3];
  }
  const uint32_t j1 = min(npts, (c + 1) * kAlignChunk);
  for (uint32_t j = c * kAlignChunk + threadIdx.x; j < j1; j += kThreads) {
    const P3 p = p3(px[j], py[j], pz[j]);
    const P3 qq = p3(qx[j], qy[j], qz[j]);
    const P3 n = p3(nx[j], ny[j], nz[j]);
    float J[NP];
    float r;
    if (KIND == 0) {
      const P3 y = p3(R0[0] * p.x + R0[1] * p.y + R0[2] * p.z + t0[0],
                      R0[3] * p.x + R0[4] * p.y + R0[5] * p.z + t0[1],
                      R0[6] * p.x + R0[7] * p.y + R0[8] * p.z + t0[2]);
      const P3 ky = cross3(w, y), kky = cross3(w, ky);
      const P3 T = add(add(add(y, mul(ky, A)), mul(kky, B)), v);
      r = dot3(n, sub(T, qq));
      // dr/dw = A (y x n) + B ((Ky) x n + y x (n x w)) + dA (n.Ky) w + dB (n.K^2 y) w
      const P3 g1 = cross3(y, n), g2 = add(cross3(ky, n), cross3(y, cross3(n, w)));
      const float s1 = dA * dot3(n, ky), s2 = dB * dot3(n, kky);
      J[0] = A * g1.x + B * g2.x + (s1 + s2) * w.x;
      J[1] = A * g1.y + B * g2.y + (s1 + s2) * w.y;
      J[2] = A * g1.z + B * g2.z + (s1 + s2) * w.z;
      J[3] = n.x, J[4] = n.y, J[5] = n.z;  // dr/dv
    } else {
      const float s = dot3(u, p) + d;
      const P3 T = sub(p, mul(u, 2.f * s));
      r = dot3(n, sub(T, qq));
      // dr/du = -2 ((n.u) p + s n); dr/dm = (I - u u^T) dr/du / |m|; dr/dd = -2 n.u
      const float nu = dot3(n, u);
      const P3 gu = mul(add(mul(p, nu), mul(n, s)), -2.f);
      const P3 gm = mul(sub(gu, mul(u, dot3(u, gu))), inv_m);
      J[0] = gm.x, J[1] = gm.y, J[2] = gm.z, J[3] = -2.f * nu;
    }
    acc[0] = __builtin_fmaf(r, r, acc[0]);
    for (int k = 0; k < NP; ++k)
      acc[1 + k] = __builtin_fmaf(2.f * r, J[k], acc[1 + k]);
    int h = 1 + NP;
    for (int a = 0; a < NP; ++a)
      for (int b = a; b < NP; ++b) {
        acc[h] = __builtin_fmaf(2.f * J[a], J[b], acc[h]);
        ++h;
      }
  }
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0) {
    float *out = part + static_cast<uint64_t>(blockIdx.x) * NV;
    for (int k = 0; k < NV; ++k)
      out[k] = acc[k];
  }
}

// One wave per mode: lane k < NV adds term k of the mode's chunk partials in
// chunk order and writes it to E, g or both mirrored entries of H.
template <int KIND>
__global__ void __launch_bounds__(kWave)
    k_align_sum(uint32_t modes, uint32_t chunks, const float *part, float *energy, float *grad,
                float *hess) {
  constexpr int NP = AlignSizes<KIND>::NP, NV = AlignSizes<KIND>::NV;
  const uint32_t m = blockIdx.x;
  const int k = threadIdx.x;
  if (m >= modes || k >= NV)
    return;
  const float *pm = part + static_cast<uint64_t>(m) * chunks * NV + k;
  float v = 0.f;
  for (uint32_t c = 0; c < chunks; ++c)
    v += pm[static_cast<uint64_t>(c) * NV];
  if (k == 0) {
    energy[m] = v;
  } else if (k <= NP) {
    grad[static_cast<uint64_t>(m) * NP + k - 1] = v;
  } else if (hess) {
    // term k is H[a][b], a <= b, in row order of the upper triangle
    int h = k - 1 - NP, a = 0;
    while (h >= NP - a)
      h -= NP - a, ++a;
    const int b = a + h;
    hess[(static_cast<uint64_t>(m) * NP + a) * NP + b] = v;
    hess[(static_cast<uint64_t>(m) * NP + b) * NP + a] = v;
  }
}

// ------------------------------------------------------------- BVH build
struct BuildNode {
  float lo[3], hi[3];
  int left = -1, right = -1, begin = 0, end = 0;
};

}  // namespace

// Host median-split build over face centroids (the widest box axis, at most
// kLeaf faces per leaf), in centre-relative FP32.
static void bvh_build(Session *s, uint64_t nv, const double *x, const double *y, const double *z,
               uint64_t nf, const int32_t *f0, const int32_t *f1, const int32_t *f2,
               const double *center) {
  if (nf == 0 || nf > kMaxFaces)
    raise(MVS_EINVAL, "a BVH needs between 1 and 2^27 faces");
  for (uint64_t f = 0; f < nf; ++f)
    for (int32_t v : {f0[f], f1[f], f2[f]})
      if (v < 0 || static_cast<uint64_t>(v) >= nv)
        raise(MVS_EINVAL, "face vertex index out of range");
  std::vector<float> vx(nv), vy(nv), vz(nv);
  for (uint64_t i = 0; i < nv; ++i) {
    vx[i] = static_cast<float>(x[i] - center[0]);
    vy[i] = static_cast<float>(y[i] - center[1]);
    vz[i] = static_cast<float>(z[i] - center[2]);
  }
  std::vector<int> ids(nf);
  std::iota(ids.begin(), ids.end(), 0);
  std::vector<float> cen(3 * nf);
  for (uint64_t f = 0; f < nf; ++f) {
    const int32_t v[3] = {f0[f], f1[f], f2[f]};
    cen[3 * f] = (vx[v[0]] + vx[v[1]] + vx[v[2]]) / 3.f;
    cen[3 * f + 1] = (vy[v[0]] + vy[v[1]] + vy[v[2]]) / 3.f;
    cen[3 * f + 2] = (vz[v[0]] + vz[v[1]] + vz[v[2]]) / 3.f;
  }
  std::vector<BuildNode> nodes;
  nodes.reserve(2 * (nf / kLeaf + 1));
  // iterative build: (node index, begin, end)
  struct Job {
    int node, begin, end, level;
  };
  std::vector<Job> jobs;
  nodes.emplace_back();
  jobs.push_back({0, 0, static_cast<int>(nf), 0});
  int depth = 0;
  while (!jobs.empty()) {
    Job jb = jobs.back();
    jobs.pop_back();
    depth = std::max(depth, jb.level);
    BuildNode nd;
    for (int k = 0; k < 3; ++k)
      nd.lo[k] = INFINITY, nd.hi[k] = -INFINITY;
    for (int i = jb.begin; i < jb.end; ++i) {
      const int f = ids[i];
      for (int32_t v : {f0[f], f1[f], f2[f]}) {
        const float c[3] = {vx[v], vy[v], vz[v]};
        for (int k = 0; k < 3; ++k)
          nd.lo[k] = std::min(nd.lo[k], c[k]), nd.hi[k] = std::max(nd.hi[k], c[k]);
      }
    }
    nd.begin = jb.begin;
    nd.end = jb.end;
    if (jb.end - jb.begin > kLeaf) {
      int axis = 0;
      for (int k = 1; k < 3; ++k)
        if (nd.hi[k] - nd.lo[k] > nd.hi[axis] - nd.lo[axis])
          axis = k;
      const int mid = (jb.begin + jb.end) / 2;
      std::nth_element(ids.begin() + jb.begin, ids.begin() + mid, ids.begin() + jb.end,
                       [&](int a, int b) {
                         return cen[3 * a + axis] < cen[3 * b + axis] ||
                                (cen[3 * a + axis] == cen[3 * b + axis] && a < b);
                       });
      nd.left = static_cast<int>(nodes.size());
      nd.right = nd.left + 1;
      nodes.emplace_back();
      nodes.emplace_back();
      jobs.push_back({nd.right, mid, jb.end, jb.level + 1});
      jobs.push_back({nd.left, jb.begin, mid, jb.level + 1});
    }
    nodes[jb.node] = nd;
  }
  // inner records numbered in build order (the root first); a leaf is a code
  std::vector<int> inner_of(nodes.size(), -1);
  int ninner = 0;
  for (size_t i = 0; i < nodes.size(); ++i)
    if (nodes[i].left >= 0)
      inner_of[i] = ninner++;
  auto ref = [&](int i) {
    const BuildNode &b = nodes[i];
    return b.left >= 0 ? inner_of[i] : -1 - (b.begin * 16 + (b.end - b.begin));
  };
  std::vector<Inner> dn(std::max(ninner, 1));
  std::memset(dn.data(), 0, dn.size() * sizeof(Inner));
  for (size_t i = 0; i < nodes.size(); ++i) {
    const BuildNode &b = nodes[i];
    if (b.left < 0)
      continue;
    Inner &r = dn[inner_of[i]];
    const BuildNode &l = nodes[b.left], &h = nodes[b.right];
    for (int k = 0; k < 3; ++k)
      r.lo0[k] = l.lo[k], r.hi0[k] = l.hi[k], r.lo1[k] = h.lo[k], r.hi1[k] = h.hi[k];
    r.ref0 = ref(b.left);
    r.ref1 = ref(b.right);
  }
  std::vector<Face> tri(nf);
  std::vector<int32_t> fid(nf), leaf_of(nf);
  for (uint64_t k = 0; k < nf; ++k) {
    const int f = ids[k];
    fid[k] = f;
    leaf_of[f] = static_cast<int32_t>(k);
    const int32_t v[3] = {f0[f], f1[f], f2[f]};
    float *corner[3] = {tri[k].a, tri[k].b, tri[k].c};
    for (int c = 0; c < 3; ++c)
      corner[c][0] = vx[v[c]], corner[c][1] = vy[v[c]], corner[c][2] = vz[v[c]];
    // unit face normal from the FP64 coordinates
    const double ax = x[v[1]] - x[v[0]], ay = y[v[1]] - y[v[0]], az = z[v[1]] - z[v[0]];
    const double bx = x[v[2]] - x[v[0]], by = y[v[2]] - y[v[0]], bz = z[v[2]] - z[v[0]];
    double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const double l = std::sqrt(cx * cx + cy * cy + cz * cz);
    if (l > 0)
      cx /= l, cy /= l, cz /= l;
    tri[k].n[0] = static_cast<float>(cx), tri[k].n[1] = static_cast<float>(cy),
    tri[k].n[2] = static_cast<float>(cz);
  }
  s->bvh_nodes.alloc(dn.size() * sizeof(Inner));
  s->bvh_tri.alloc(tri.size() * sizeof(Face));
  s->bvh_fid.alloc(fid.size() * sizeof(int32_t));
  s->bvh_leaf_of.alloc(leaf_of.size() * sizeof(int32_t));
  MVS_HIP(hipMemcpy(s->bvh_leaf_of.ptr, leaf_of.data(), leaf_of.size() * sizeof(int32_t),
                    hipMemcpyHostToDevice));
  MVS_HIP(hipMemcpy(s->bvh_nodes.ptr, dn.data(), dn.size() * sizeof(Inner),
                    hipMemcpyHostToDevice));
  MVS_HIP(hipMemcpy(s->bvh_tri.ptr, tri.data(), tri.size() * sizeof(Face),
                    hipMemcpyHostToDevice));
  MVS_HIP(hipMemcpy(s->bvh_fid.ptr, fid.data(), fid.size() * sizeof(int32_t),
                    hipMemcpyHostToDevice));
  s->bvh_faces = nf;
  s->bvh_root = ref(0);
  // one waiting sibling per level plus the two children just pushed
  s->bvh_stack = depth + 1;
}

// Dynamic LDS of a closest-point launch: the stack of every thread.
static size_t stack_lds(const Session *s) {
  return static_cast<size_t>(s->bvh_stack) * kThreads * (sizeof(int) + sizeof(float));
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_bvh_build(mvs_session *sp, uint64_t nv, const double *x, const double *y,
                  const double *z, uint64_t nf, const int32_t *f0, const int32_t *f1,
                  const int32_t *f2, const double center[3]) {
  return guarded([&] {
    if (!x || !y || !z || !f0 || !f1 || !f2 || !center)
      raise(MVS_EINVAL, "null argument");
    bvh_build(as_session(sp), nv, x, y, z, nf, f0, f1, f2, center);
  });
}

int mvs_dev_closest_points(mvs_session *sp, uint64_t n, const float *pts, float *q, float *nrm,
                           float *dist2, int32_t *face) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!s->bvh_faces)
      raise(MVS_EINVAL, "closest points need a BVH (mvs_bvh_build)");
    if (n && (!pts || !q || !nrm || !dist2))
      raise(MVS_EINVAL, "null argument");
    if (n)
      MVS_LAUNCH(s, "closest", k_closest, blocks_for(n, kThreads), kThreads, stack_lds(s),
                 reinterpret_cast<const Inner *>(s->bvh_nodes.ptr),
                 reinterpret_cast<const Face *>(s->bvh_tri.ptr), s->bvh_root, s->bvh_stack,
                 reinterpret_cast<const int32_t *>(s->bvh_fid.ptr), n, pts, q, nrm, dist2, face);
  });
}

int mvs_dev_closest_modes(mvs_session *sp, int kind, uint32_t modes, uint32_t npts,
                          const float *theta, const float *base, const float *pts,
                          int shared_pts, float *q, float *nrm, float *dist2, int32_t *face) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!s->bvh_faces)
      raise(MVS_EINVAL, "closest points need a BVH (mvs_bvh_build)");
    if (kind != 0 && kind != 1)
      raise(MVS_EINVAL, "kind is 0 (rigid) or 1 (mirror)");
    if (kind == 0 && !base)
      raise(MVS_EINVAL, "rigid modes need their base transforms");
    if (!theta || !pts || !q || !nrm || !dist2)
      raise(MVS_EINVAL, "null argument");
    if (modes == 0 || npts == 0)
      return;
    const unsigned per = blocks_for(npts, kThreads);
    if (static_cast<uint64_t>(per) * modes > 0x7fffffffULL)
      raise(MVS_EOVERFLOW, "too many closest-point queries for one launch");
    const uint64_t stride = shared_pts ? 0 : 3ULL * npts;
    const Inner *inner = reinterpret_cast<const Inner *>(s->bvh_nodes.ptr);
    const Face *faces = reinterpret_cast<const Face *>(s->bvh_tri.ptr);
    const int root = s->bvh_root, depth = s->bvh_stack;
    const size_t lds = stack_lds(s);
    const int32_t *fid = reinterpret_cast<const int32_t *>(s->bvh_fid.ptr);
    const int32_t *leaf_of = reinterpret_cast<const int32_t *>(s->bvh_leaf_of.ptr);
    const uint32_t nf = static_cast<uint32_t>(s->bvh_faces);
    if (kind == 0)
      MVS_LAUNCH(s, "closest", k_closest_modes<0>, per * modes, kThreads, lds, inner, faces,
                 root, depth, modes, npts, per, theta, base, pts, stride, q, nrm, dist2, face, fid, leaf_of,
                 nf);
    else
      MVS_LAUNCH(s, "closest", k_closest_modes<1>, per * modes, kThreads, lds, inner, faces,
                 root, depth, modes, npts, per, theta, base, pts, stride, q, nrm, dist2, face, fid, leaf_of,
                 nf);
  });
}

int mvs_dev_align_energy(mvs_session *sp, int kind, uint32_t modes, uint32_t npts,
                         const float *theta, const float *base, const float *pts, const float *q,
                         const float *nrm, float *energy, float *grad, float *hess) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (kind != 0 && kind != 1)
      raise(MVS_EINVAL, "kind is 0 (rigid) or 1 (mirror)");
    if (kind == 0 && !base)
      raise(MVS_EINVAL, "rigid modes need their base transforms");
    if (!theta || !pts || !q || !nrm || !energy || !grad)
      raise(MVS_EINVAL, "null argument");
    if (modes == 0)
      return;
    const uint32_t chunks = std::max<uint32_t>(1, (npts + kAlignChunk - 1) / kAlignChunk);
    if (static_cast<uint64_t>(modes) * chunks > 0x7fffffffULL)
      raise(MVS_EOVERFLOW, "too many energy blocks for one launch");
    const int nv = kind == 0 ? AlignSizes<0>::NV : AlignSizes<1>::NV;
    s->align_part.ensure(static_cast<size_t>(modes) * chunks * nv);
    float *part = s->align_part.ptr;
    if (kind == 0) {
      MVS_LAUNCH(s, "align", k_align<0>, modes * chunks, kThreads, 0, modes, npts, chunks,
                 theta, base, pts, q, nrm, part);
      MVS_LAUNCH(s, "align_sum", k_align_sum<0>, modes, kWave, 0, modes, chunks, part, energy,
                 grad, hess);
    } else {
      MVS_LAUNCH(s, "align", k_align<1>, modes * chunks, kThreads, 0, modes, npts, chunks,
                 theta, base, pts, q, nrm, part);
      MVS_LAUNCH(s, "align_sum", k_align_sum<1>, modes, kWave, 0, modes, chunks, part, energy,
                 grad, hess);
    }
  });
}

}  // extern "C"
