// Vote emission: one optional mirror vote and two rigid votes per pair
// (restates src/cpu/transformations.cpp:17-73).
//
// Two outputs share the pair geometry:
//   * compact records on the device (the fast path): rigid votes in two
//     slots per pair (q[4] with w >= 0, translation about the model centre
//     t_c[3], pair id) and mirror votes in pair order (n[3], offset about the
//     centre d_c, pair id), FP64 or FP32 SoA streams;
//   * the reference's 28-lane Votes layout in global vote order (the
//     Backend::transformations drop-in).
// Mirror votes need an exclusive prefix over pairs; it is computed per
// wavefront of 64 consecutive pairs (vote count pass), scanned, and each
// emitting wave adds its lane prefix from one ballot -- no block barriers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>

#include "pairing.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr unsigned kXcds = 8;  // MI355X: 8 dies, blocks dealt to them round-robin

// Wave w of a pair pass, with blocks remapped so that each die takes a
// contiguous run of 64-pair chunks: consecutive pairs share their sources,
// whose records then stay in that die's L2.  Launch grids are rounded up to
// a multiple of kXcds (waves past the pairs return at once).
__device__ inline uint64_t pair_wave(uint64_t P) {
  const uint64_t nw = (P + kWave - 1) / kWave;
  const unsigned nb = static_cast<unsigned>((nw + kThreads / kWave - 1) / (kThreads / kWave));
  const unsigned per = (nb + kXcds - 1) / kXcds;
  const unsigned logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  return static_cast<uint64_t>(logical) * (kThreads / kWave) + threadIdx.x / kWave;
}
inline unsigned pair_grid(uint64_t nw) {
  const unsigned nb = blocks_for(nw, kThreads / kWave);
  return ((nb + kXcds - 1) / kXcds) * kXcds;
}

// Pair geometry of transformations.cpp:34-47 in FP64.
struct Geo64 {
  V3 p, q, an, ad, ax, bn, bd, bx, n;
  double d;
  M3 mirror;
  bool refl;
};
__device__ inline Geo64 geo64(V3 p, V3 q, V3 an, V3 ad, V3 ax, V3 bn, V3 bd, V3 bx) {
  Geo64 g;
  g.p = p, g.q = q, g.an = an, g.ad = ad, g.ax = ax, g.bn = bn, g.bd = bd, g.bx = bx;
  V3 n = unit(q - p);
  int dom = 0;
  for (int k = 1; k < 3; ++k)
    if (fabs(comp(n, k)) > fabs(comp(n, dom)))
      dom = k;
  if (comp(n, dom) < 0)
    n = n * -1;
  g.n = n;
  g.d = -dot(n, vdivs(p + q, 2));
  g.mirror = reflection(n);
  g.refl = dot(apply(g.mirror, an), bn) > 0.75 && fabs(dot(apply(g.mirror, ad), bd)) > 0.75;
  return g;
}
__device__ inline M3 rigid64(const Geo64 &g, double sign) {
  return mul_bt(frame(g.bd * sign, g.bx * sign, g.bn), frame(g.ad, g.ax, g.an));
}

// Rotation matrix -> unit quaternion (w, x, y, z) with w >= 0 (Shepperd's
// method: the largest of the four diagonal combinations gives the pivot
// component, the off-diagonal sums and differences the other three).  The
// pivot is chosen by selects, not branches, so the lanes of a wave that pick
// different pivots do not serialise; one root and one reciprocal per matrix.
__device__ inline float quat_inv(float s) { return __builtin_amdgcn_rcpf(s); }
__device__ inline double quat_inv(double s) { return 1.0 / s; }
template <class R> __device__ inline void quat(const R *m, R *q) {
  const R tr = m[0] + m[4] + m[8];
  const bool c0 = tr > R(0);
  const bool c1 = !c0 && m[0] > m[4] && m[0] > m[8];
  const bool c2 = !c0 && !c1 && m[4] > m[8];
  const bool c3 = !c0 && !c1 && !c2;
  const R d = c0 ? tr + R(1)
                 : (c1 ? R(1) + m[0] - m[4] - m[8]
                       : (c2 ? R(1) + m[4] - m[0] - m[8] : R(1) + m[8] - m[0] - m[4]));
  const R s = sqrt(d) * R(2), piv = R(0.25) * s, inv = quat_inv(s);
  const R a = (m[7] - m[5]) * inv, b = (m[2] - m[6]) * inv, c = (m[3] - m[1]) * inv;
  const R e = (m[1] + m[3]) * inv, f = (m[2] + m[6]) * inv, g = (m[5] + m[7]) * inv;
  R w = c0 ? piv : (c1 ? a : (c2 ? b : c));
  R x = c0 ? a : (c1 ? piv : (c2 ? e : f));
  R y = c0 ? b : (c1 ? e : (c2 ? piv : g));
  R z = c0 ? c : (c1 ? f : (c2 ? g : piv));
  (void)c3;
  if (w < R(0))
    w = -w, x = -x, y = -y, z = -z;
  q[0] = w, q[1] = x, q[2] = y, q[3] = z;
}

struct F3 {
  float x, y, z;
};
__device__ inline F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float fcomp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Three consecutive fields of sample i's record (AoS, kRec scalars each).
__device__ inline V3 field3(const double *r, uint64_t S, int l, uint64_t i) {
  (void)S;
  const double *p = r + i * kRec + l;
  return v3(p[0], p[1], p[2]);
}
__device__ inline F3 field3f(const float *r, uint64_t S, int l, uint64_t i) {
  (void)S;
  const float *p = r + i * kRec + l;
  return f3(p[0], p[1], p[2]);
}
// A whole FP32 sample record (one 64-byte line) in registers: four 16-byte
// loads issued together, so a pair's two records arrive in one round trip.
struct Rec32 {
  float4 v[kRec / 4];
};
__device__ inline Rec32 load_rec32(const float *s32, uint64_t i) {
  const float4 *p = reinterpret_cast<const float4 *>(s32 + i * kRec);
  Rec32 r;
  for (int k = 0; k < kRec / 4; ++k)
    r.v[k] = p[k];
  return r;
}
__device__ inline float rget(const Rec32 &r, int l) {
  const float4 &q = r.v[l >> 2];
  const int c = l & 3;
  return c == 0 ? q.x : (c == 1 ? q.y : (c == 2 ? q.z : q.w));
}
__device__ inline F3 rf3(const Rec32 &r, int l) { return f3(rget(r, l), rget(r, l + 1), rget(r, l + 2)); }

__device__ inline Geo64 geo_samples(const double *s, uint64_t S, uint64_t i, uint64_t o) {
  return geo64(field3(s, S, L_X, i), field3(s, S, L_X, o), field3(s, S, L_NX, i),
               field3(s, S, L_DX, i), field3(s, S, L_XX, i), field3(s, S, L_NX, o),
               field3(s, S, L_DX, o), field3(s, S, L_XX, o));
}

struct VoteArgs {
  const double *s64;
  const float *s32;
  uint64_t S;
  const int32_t *a, *b;  // pair ordinals
  uint64_t P;
  double cx, cy, cz;
  float grefl, gdom;  // guard bands of the FP32 path (set per launch)
  float g0, g1;       // band = g0 + g1 * (largest FP32 coordinate)
  const unsigned *rmax_bits;
  uint32_t *wave_rr;  // per wave: FP64 rechecks of the mirror tests
  unsigned *bad;      // set when a pair ordinal is outside [0, S)
  Segs sg;            // batches: the centre is the pair's mesh centre
};
// Model centre of the mesh that sample i belongs to.
__device__ inline V3 centre_of(const VoteArgs &A, uint64_t i) {
  if (!A.sg.seg)
    return v3(A.cx, A.cy, A.cz);
  const uint64_t m = A.sg.seg[i];
  return v3(A.sg.center[3 * m], A.sg.center[3 * m + 1], A.sg.center[3 * m + 2]);
}
// Both ordinals of pair k index the sample records.
__device__ inline bool pair_in_range(const int32_t *a, const int32_t *b, uint64_t k, uint64_t S) {
  int32_t i = a[k], o = b[k];
  return i >= 0 && o >= 0 && static_cast<uint64_t>(i) < S && static_cast<uint64_t>(o) < S;
}
template <bool kGuarded> __device__ inline VoteArgs with_bands(const VoteArgs &in) {
  VoteArgs A = in;
  if (kGuarded) {
    float g = in.g0 + in.g1 * __uint_as_float(*in.rmax_bits);
    A.grefl = g;
    A.gdom = g;
  }
  return A;
}

// Mirror decision for pair k.  The FP32 result stands unless a test value is
// inside its guard band or the direction is degenerate; then FP64 decides.
// (The decision does not depend on the sign of n, so axis ties do not matter.)
// ra, rb: the FP32 records of samples i and o (read only when kGuarded).
template <bool kGuarded>
__device__ inline bool pair_reflection(const VoteArgs &A, uint64_t i, uint64_t o,
                                       const Rec32 &ra, const Rec32 &rb, unsigned &rr) {
  const uint64_t S = A.S;
  if (kGuarded) {
    F3 p = rf3(ra, L_X), q = rf3(rb, L_X);
    F3 dv = q - p;
    float len = sqrtf(fdot(dv, dv));
    if (len > 1e-6f) {
      F3 n = dv * (1.0f / len);
      F3 an = rf3(ra, L_NX), ad = rf3(ra, L_DX);
      F3 bn = rf3(rb, L_NX), bd = rf3(rb, L_DX);
      // (I - 2 n n^T) v = v - 2 (n.v) n
      F3 ma = an - n * (2.0f * fdot(n, an)), md = ad - n * (2.0f * fdot(n, ad));
      float r1 = fdot(ma, bn), r2 = fabsf(fdot(md, bd));
      if (r1 < 0.75f - A.grefl || r2 < 0.75f - A.grefl)
        return false;
      if (r1 > 0.75f + A.grefl && r2 > 0.75f + A.grefl)
        return true;
    }
    ++rr;
  }
  return geo_samples(A.s64, S, i, o).refl;
}

template <bool kGuarded>
__global__ void __launch_bounds__(kThreads) k_vote_count(VoteArgs Ain, uint32_t *wave_refl) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w = pair_wave(Ain.P);
  const uint64_t k = w * kWave + lane;
  if (w * kWave >= Ain.P)
    return;
  const VoteArgs A = with_bands<kGuarded>(Ain);
  unsigned rr = 0;
  const bool inp = k < A.P;
  const int32_t ia = inp ? A.a[k] : 0, ib = inp ? A.b[k] : 0;
  const bool in = inp && ia >= 0 && ib >= 0 && static_cast<uint64_t>(ia) < A.S &&
                  static_cast<uint64_t>(ib) < A.S;
  if (inp && !in)
    atomicOr(A.bad, 1u);
  const uint64_t i = in ? static_cast<uint64_t>(ia) : 0, o = in ? static_cast<uint64_t>(ib) : 0;
  Rec32 ra{}, rb{};
  if (kGuarded) {  // sample 0 stands in for dead lanes: every load is in range
    ra = load_rec32(A.s32, i);
    rb = load_rec32(A.s32, o);
  }
  bool refl = in && pair_reflection<kGuarded>(A, i, o, ra, rb, rr);
  uint64_t m = ballot(refl);
  rr = wave_sum(rr);
  if (lane == 0) {
    wave_refl[w] = static_cast<uint32_t>(__popcll(m));
    A.wave_rr[w] = rr;
  }
}

template <class R> struct VoteOut {
  R *q4;        // rigid: 4 lanes, stride 2P
  R *t3;        // rigid: 3 lanes, stride 2P
  int32_t *rp;  // rigid pair ids (null: not written, the pair of vote v is v / 2)
  R *n3;        // mirror: 3 lanes, stride R_total
  R *d1;        // mirror offsets
  int32_t *fp;  // mirror pair ids
  uint64_t R_total;
};

template <class R> struct Vec2T;
template <> struct Vec2T<float> { using T = float2; };
template <> struct Vec2T<double> { using T = double2; };

template <bool kGuarded, class R>
__global__ void __launch_bounds__(kThreads)
    k_vote_emit(VoteArgs Ain, const uint64_t *wave_off, VoteOut<R> out) {
  using V2 = typename Vec2T<R>::T;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w = pair_wave(Ain.P);
  const uint64_t k = w * kWave + lane;
  if (w * kWave >= Ain.P)
    return;
  const VoteArgs A = with_bands<kGuarded>(Ain);
  const uint64_t S = A.S, P2 = 2 * A.P;
  // out-of-range pairs were flagged by the count pass and are skipped here;
  // the pair's ordinals and (FP32) records are read once, up front
  const bool inp = k < A.P;
  const int32_t ia = inp ? A.a[k] : 0, ib = inp ? A.b[k] : 0;
  const bool live = inp && ia >= 0 && ib >= 0 && static_cast<uint64_t>(ia) < S &&
                    static_cast<uint64_t>(ib) < S;
  const uint64_t i = live ? static_cast<uint64_t>(ia) : 0, o = live ? static_cast<uint64_t>(ib) : 0;
  Rec32 ra{}, rb{};
  if (kGuarded) {  // sample 0 stands in for dead lanes: every load is in range
    ra = load_rec32(A.s32, i);
    rb = load_rec32(A.s32, o);
  }
  unsigned rr = 0;
  bool refl = live && pair_reflection<kGuarded>(A, i, o, ra, rb, rr);
  const uint64_t slot = wave_off[w] + lanes_below(ballot(refl));
  if (!live)
    return;
  R rq[2][4], rt[2][3], pn[3] = {R(0), R(0), R(0)}, pd = R(0);
  if (!kGuarded) {
    Geo64 g = geo_samples(A.s64, S, i, o);
    const V3 c = centre_of(A, i);
    for (int sg = 0; sg < 2; ++sg) {
      M3 rot = rigid64(g, sg == 0 ? 1.0 : -1.0);
      V3 t = g.q - apply(rot, g.p);
      V3 tc = apply(rot, c) + t - c;  // translation about the model centre
      double qq[4];
      quat<double>(rot.m, qq);
      for (int kk = 0; kk < 4; ++kk)
        rq[sg][kk] = static_cast<R>(qq[kk]);
      rt[sg][0] = static_cast<R>(tc.x), rt[sg][1] = static_cast<R>(tc.y),
      rt[sg][2] = static_cast<R>(tc.z);
    }
    if (refl) {
      pn[0] = static_cast<R>(g.n.x), pn[1] = static_cast<R>(g.n.y), pn[2] = static_cast<R>(g.n.z);
      pd = static_cast<R>(g.d + dot(g.n, c));
    }
  } else {
    // FP32 frames in centre-relative coordinates: t_c = q' - R p'.  The
    // vote values (not the mirror decision, which the count pass must repeat
    // bit for bit) may fuse products into sums: one rounding fewer each.
#pragma clang fp contract(fast)
    F3 p = rf3(ra, L_X), q = rf3(rb, L_X);
    F3 an = rf3(ra, L_NX), ad = rf3(ra, L_DX), ax = rf3(ra, L_XX);
    F3 bn = rf3(rb, L_NX), bd = rf3(rb, L_DX), bx = rf3(rb, L_XX);
    // R = B A^T with B = [+-dmin_b, +-dmax_b, n_b]: the two signs share the
    // first two column products u and the third v (R+ = u + v, R- = -u + v,
    // the same roundings as forming each product in full)
    float u[9], v9[9];
    {
      const float Bc[3][3] = {{bd.x, bx.x, bn.x}, {bd.y, bx.y, bn.y}, {bd.z, bx.z, bn.z}};
      const float Ac[3][3] = {{ad.x, ax.x, an.x}, {ad.y, ax.y, an.y}, {ad.z, ax.z, an.z}};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          u[r * 3 + c] = Bc[r][0] * Ac[c][0] + Bc[r][1] * Ac[c][1];
          v9[r * 3 + c] = Bc[r][2] * Ac[c][2];
        }
    }
    for (int sg = 0; sg < 2; ++sg) {
      float m[9];
      for (int e = 0; e < 9; ++e)
        m[e] = (sg == 0 ? u[e] : -u[e]) + v9[e];
      float qq[4];
      quat<float>(m, qq);
      for (int kk = 0; kk < 4; ++kk)
        rq[sg][kk] = static_cast<R>(qq[kk]);
      rt[sg][0] = static_cast<R>(q.x - (m[0] * p.x + m[1] * p.y + m[2] * p.z));
      rt[sg][1] = static_cast<R>(q.y - (m[3] * p.x + m[4] * p.y + m[5] * p.z));
      rt[sg][2] = static_cast<R>(q.z - (m[6] * p.x + m[7] * p.y + m[8] * p.z));
    }
    if (refl) {
      F3 dv = q - p;
      float len = sqrtf(fdot(dv, dv));
      F3 n = dv * (1.0f / len);
      float ax_ = fabsf(n.x), ay_ = fabsf(n.y), az_ = fabsf(n.z);
      float m1 = fmaxf(ax_, fmaxf(ay_, az_));
      float m2 = fmaxf(fminf(ax_, ay_), fminf(fmaxf(ax_, ay_), az_));
      if (m1 - m2 <= A.gdom || !(len > 1e-6f)) {
        // near-tied dominant axis: the sign of n comes from FP64
        Geo64 g = geo_samples(A.s64, S, i, o);
        const V3 c = centre_of(A, i);
        pn[0] = static_cast<R>(g.n.x), pn[1] = static_cast<R>(g.n.y),
        pn[2] = static_cast<R>(g.n.z);
        pd = static_cast<R>(g.d + dot(g.n, c));
      } else {
        int dom = 0;
        for (int kk = 1; kk < 3; ++kk)
          if (fabsf(fcomp(n, kk)) > fabsf(fcomp(n, dom)))
            dom = kk;
        if (fcomp(n, dom) < 0)
          n = n * -1.0f;
        pn[0] = static_cast<R>(n.x), pn[1] = static_cast<R>(n.y), pn[2] = static_cast<R>(n.z);
        pd = static_cast<R>(-fdot(n, (p + q) * 0.5f));  // -n.(p'+q')/2 = d + n.c
      }
    }
  }
  // rigid slots 2k (sign +1) and 2k+1 (sign -1): one 2-wide store per lane
  for (int kk = 0; kk < 4; ++kk) {
    V2 v;
    v.x = rq[0][kk];
    v.y = rq[1][kk];
    *reinterpret_cast<V2 *>(out.q4 + kk * P2 + 2 * k) = v;
  }
  for (int kk = 0; kk < 3; ++kk) {
    V2 v;
    v.x = rt[0][kk];
    v.y = rt[1][kk];
    *reinterpret_cast<V2 *>(out.t3 + kk * P2 + 2 * k) = v;
  }
  if (out.rp) {  // optional: rigid vote v belongs to pair v / 2
    int2 ids;
    ids.x = ids.y = static_cast<int32_t>(k);
    *reinterpret_cast<int2 *>(out.rp + 2 * k) = ids;
  }
  if (refl && slot < out.R_total) {
    const uint64_t RT = out.R_total;
    out.n3[slot] = pn[0];
    out.n3[RT + slot] = pn[1];
    out.n3[2 * RT + slot] = pn[2];
    out.d1[slot] = pd;
    out.fp[slot] = static_cast<int32_t>(k);
  }
}

// --------------------------------------- reference-layout transformations
struct TransArgs {
  const double *x, *y, *z, *feat;
  uint64_t n;
  const int32_t *a, *b;
  uint64_t P;
  double box, cx, cy, cz;
};
__device__ inline Geo64 geo_mesh(const TransArgs &A, uint64_t k) {
  uint64_t i = static_cast<uint64_t>(A.a[k]), o = static_cast<uint64_t>(A.b[k]);
  const double *f = A.feat;
  const uint64_t n = A.n;
  return geo64(v3(A.x[i], A.y[i], A.z[i]), v3(A.x[o], A.y[o], A.z[o]),
               v3(f[i], f[n + i], f[2 * n + i]), v3(f[3 * n + i], f[4 * n + i], f[5 * n + i]),
               v3(f[6 * n + i], f[7 * n + i], f[8 * n + i]), v3(f[o], f[n + o], f[2 * n + o]),
               v3(f[3 * n + o], f[4 * n + o], f[5 * n + o]),
               v3(f[6 * n + o], f[7 * n + o], f[8 * n + o]));
}
__global__ void k_trans_count(TransArgs A, uint32_t *cnt) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < A.P)
    cnt[k] = geo_mesh(A, k).refl ? 3u : 2u;
}
__device__ inline void put28(double *lanes, uint64_t V, uint64_t row, const M3 &r, V3 t, V3 n,
                             double d, const double *desc) {
  for (int k = 0; k < 9; ++k)
    lanes[k * V + row] = r.m[k];
  double v[7] = {t.x, t.y, t.z, n.x, n.y, n.z, d};
  for (int k = 0; k < 7; ++k)
    lanes[(9 + k) * V + row] = v[k];
  for (int k = 0; k < 12; ++k)
    lanes[(16 + k) * V + row] = desc[k];
}
__global__ void k_trans_fill(TransArgs A, const uint64_t *off, uint64_t V, double *lanes,
                             int32_t *pair, uint8_t *kind) {
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= A.P)
    return;
  Geo64 g = geo_mesh(A, k);
  uint64_t row = off[k];
  V3 c = v3(A.cx, A.cy, A.cz);
  if (g.refl) {
    double desc[12] = {g.n.x, g.n.y, g.n.z, (g.d + dot(g.n, c)) / A.box, 0, 0, 0, 0, 0, 0, 0, 0};
    put28(lanes, V, row, g.mirror, g.n * (-2 * g.d), g.n, g.d, desc);
    pair[row] = static_cast<int32_t>(k);
    kind[row] = 1;
    ++row;
  }
  for (int sg = 0; sg < 2; ++sg) {
    M3 rot = rigid64(g, sg == 0 ? 1.0 : -1.0);
    V3 t = g.q - apply(rot, g.p);
    V3 cen = vdivs(apply(rot, c) + t - c, A.box);
    double desc[12];
    for (int kk = 0; kk < 9; ++kk)
      desc[kk] = rot.m[kk] / 2;
    desc[9] = cen.x, desc[10] = cen.y, desc[11] = cen.z;
    put28(lanes, V, row, rot, t, v3(0, 0, 0), 0, desc);
    pair[row] = static_cast<int32_t>(k);
    kind[row] = 0;
    ++row;
  }
}

}  // namespace

static VoteArgs vote_args(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                          const double *center, bool guarded, const Segs &segs) {
  PairPlan &plan = s->pair_plan;
  if (!plan.valid)
    raise(MVS_EINVAL, "vote emission needs the sample plan of a pair count");
  if (guarded && plan.precision != MVS_FP32_GUARDED)
    raise(MVS_EINVAL, "FP32-guarded votes need an FP32-guarded pair plan");
  VoteArgs A{};
  A.s64 = plan_samples64(s);
  A.s32 = guarded ? plan_samples32(s) : nullptr;
  A.S = plan.S;
  A.a = a;
  A.b = b;
  A.P = P;
  A.cx = center ? center[0] : 0;
  A.cy = center ? center[1] : 0;
  A.cz = center ? center[2] : 0;
  // Every emitted pair is >= dist = 0.03 box apart and every FP32
  // coordinate is bounded by rmax (measured in the gather), so the FP32
  // direction carries a relative error of at most ~8 eps32 rmax / dist; the
  // dot tests and the axis tie see at most a few times that.  Band:
  // 2e-4 + 64 eps32 rmax / dist.
  A.grefl = A.gdom = 0.f;
  A.g0 = 2e-4f;
  A.g1 = static_cast<float>(64.0 * 5.960464477539063e-08 / (0.03 * plan.box));
  A.rmax_bits = plan_flags(s) + kRmaxWord;
  const uint64_t nw = (P + kWave - 1) / kWave;
  A.wave_rr = scratch_as<uint32_t>(s, kSlotRefl, nw ? nw : 1);
  A.bad = plan_flags(s) + kVoteBadWord;
  A.sg = segs;
  return A;
}

void vote_offsets_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                         int precision, uint64_t *total_dev, const Segs &segs,
                         bool bad_cleared) {
  if (precision != MVS_FP64 && precision != MVS_FP32_GUARDED)
    raise(MVS_EINVAL, "unknown precision");
  const bool guarded = precision == MVS_FP32_GUARDED;
  VoteArgs A = vote_args(s, P, a, b, nullptr, guarded, segs);
  const uint64_t nw = (P + kWave - 1) / kWave;
  uint32_t *wr = scratch_as<uint32_t>(s, kSlotVoteBlock, nw ? nw : 1);
  uint64_t *off = scratch_as<uint64_t>(s, kSlotVoteBlockOff, nw + 1);
  if (!bad_cleared)
    zero_device(s, A.bad, sizeof(unsigned));
  if (P) {
    unsigned grid = pair_grid(nw);
    if (guarded)
      MVS_LAUNCH(s, "vote_count", k_vote_count<true>, grid, kThreads, 0, A, wr);
    else
      MVS_LAUNCH(s, "vote_count", k_vote_count<false>, grid, kThreads, 0, A, wr);
  }
  exclusive_scan_u32(s, wr, off, nw, total_dev);
}

void vote_emit_strided(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                       const double *center, int precision, uint64_t stride, void *q4, void *t3,
                       int32_t *rp, void *n3, void *d1, int32_t *fp, const Segs &segs) {
  if (P == 0)
    return;
  const bool guarded = precision == MVS_FP32_GUARDED;
  VoteArgs A = vote_args(s, P, a, b, center, guarded, segs);
  const uint64_t nw = (P + kWave - 1) / kWave;
  uint64_t *off = scratch_as<uint64_t>(s, kSlotVoteBlockOff, nw + 1);
  unsigned grid = pair_grid(nw);
  if (guarded) {
    VoteOut<float> o{static_cast<float *>(q4), static_cast<float *>(t3), rp,
                     static_cast<float *>(n3), static_cast<float *>(d1), fp, stride};
    MVS_LAUNCH(s, "vote_emit", (k_vote_emit<true, float>), grid, kThreads, 0, A, off, o);
  } else {
    VoteOut<double> o{static_cast<double *>(q4), static_cast<double *>(t3), rp,
                      static_cast<double *>(n3), static_cast<double *>(d1), fp, stride};
    MVS_LAUNCH(s, "vote_emit", (k_vote_emit<false, double>), grid, kThreads, 0, A, off, o);
  }
}

uint64_t vote_total_checked(Session *s) {
  // words kVoteTotalWord .. kVoteTotalWord + 2: total, (unused), bad flag
  uint64_t w[3];
  read_back(s, w, reinterpret_cast<uint64_t *>(plan_flags(s)) + kVoteTotalWord, sizeof w);
  if (static_cast<uint32_t>(w[2]))
    raise(MVS_EINVAL, "pair ordinal outside the sample set");
  return w[0];
}

void vote_count_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b, int precision,
                       uint64_t *R_out) {
  uint64_t *tot = reinterpret_cast<uint64_t *>(plan_flags(s)) + kVoteTotalWord;
  vote_offsets_device(s, P, a, b, precision, tot);
  uint64_t R = vote_total_checked(s);
  s->vote_plan.P = P;
  s->vote_plan.R = R;
  s->vote_plan.precision = precision;
  s->vote_plan.valid = true;
  *R_out = R;
}

void vote_emit_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                      const double *center, int precision, void *q4, void *t3, int32_t *rp,
                      void *n3, void *d1, int32_t *fp) {
  VotePlan &vp = s->vote_plan;
  if (!vp.valid || vp.P != P || vp.precision != precision)
    raise(MVS_EINVAL, "vote emission without a matching vote count");
  vote_emit_strided(s, P, a, b, center, precision, vp.R, q4, t3, rp, n3, d1, fp);
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_vote_count(mvs_session *sp, uint64_t P, const int32_t *a, const int32_t *b,
                       int precision, uint64_t *reflections) {
  return guarded([&] {
    if (!reflections)
      raise(MVS_EINVAL, "null output pointer");
    vote_count_device(as_session(sp), P, a, b, precision, reflections);
  });
}

int mvs_dev_vote_emit(mvs_session *sp, uint64_t P, const int32_t *a, const int32_t *b, double box,
                      const double center[3], int precision, void *q4, void *t3, int32_t *rp,
                      void *n3, void *d1, int32_t *fp) {
  return guarded([&] {
    (void)box;
    if (!center)
      raise(MVS_EINVAL, "null centre");
    vote_emit_device(as_session(sp), P, a, b, center, precision, q4, t3, rp, n3, d1, fp);
  });
}

int mvs_transformations(mvs_session *sp, const mvs_points *v, const mvs_features *f,
                        uint64_t P, const int32_t *a, const int32_t *b, double box,
                        const double center[3], uint64_t max_votes, uint64_t *count) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!v || !f || !count || (P && (!a || !b)) || !center)
      raise(MVS_EINVAL, "null argument");
    *count = 0;
    s->out_count = 0;
    if (P == 0)
      return;
    uint64_t n = v->count;
    if (f->count != n)
      raise(MVS_EINVAL, "SoA component lengths differ");
    if (!(box > 0) || !std::isfinite(box) || !std::isfinite(center[0]) ||
        !std::isfinite(center[1]) || !std::isfinite(center[2]))
      raise(MVS_EINVAL, "transformations require a positive finite box size and finite center");
    if (P > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
      raise(MVS_EOVERFLOW, "pair ids exceed int32 range");
    for (uint64_t i = 0; i < P; ++i)
      if (a[i] < 0 || b[i] < 0 || static_cast<uint64_t>(a[i]) >= n ||
          static_cast<uint64_t>(b[i]) >= n)
        raise(MVS_EINVAL, "pair vertex index out of range");
    DBuf<double> dv(3 * n), feat(11 * n);
    DBuf<int32_t> da(P), db(P);
    DBuf<uint32_t> cnt(P);
    DBuf<uint64_t> off(P + 1);
    copy_h2d(s, dv.ptr, v->x, n * 8);
    copy_h2d(s, dv.ptr + n, v->y, n * 8);
    copy_h2d(s, dv.ptr + 2 * n, v->z, n * 8);
    for (int k = 0; k < 11; ++k)
      copy_h2d(s, feat.ptr + k * n, f->lane[k], n * 8);
    copy_h2d(s, da.ptr, a, P * 4);
    copy_h2d(s, db.ptr, b, P * 4);
    TransArgs A{dv.ptr, dv.ptr + n, dv.ptr + 2 * n, feat.ptr, n, da.ptr, db.ptr, P,
                box,    center[0],  center[1],     center[2]};
    MVS_LAUNCH(s, "trans_count", k_trans_count, blocks_for(P, 256), 256, 0, A, cnt.ptr);
    exclusive_scan_u32(s, cnt.ptr, off.ptr, P);
    uint64_t V = 0;
    copy_d2h(s, &V, off.ptr + P, sizeof V);
    if (V > max_votes || V > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
      raise(MVS_EBUDGET, "Vote budget exceeded; lower --sample-ratio or tighten --pair-threshold");
    s->out_lanes.ensure(28 * V);
    s->out_pair.ensure(V);
    s->out_kind.ensure(V);
    MVS_LAUNCH(s, "trans_fill", k_trans_fill, blocks_for(P, 256), 256, 0, A, off.ptr, V,
               s->out_lanes.ptr, s->out_pair.ptr, s->out_kind.ptr);
    sync(s);
    s->out_count = V;
    *count = V;
  });
}

int mvs_transformations_fetch(mvs_session *sp, int32_t *pair, uint8_t *kind,
                              double *const lanes[28]) {
  return guarded([&] {
    Session *s = as_session(sp);
    uint64_t V = s->out_count;
    copy_d2h(s, pair, s->out_pair.ptr, V * 4);
    copy_d2h(s, kind, s->out_kind.ptr, V);
    for (int k = 0; k < 28; ++k)
      copy_d2h(s, lanes[k], s->out_lanes.ptr + k * V, V * 8);
  });
}

}  // extern "C"
