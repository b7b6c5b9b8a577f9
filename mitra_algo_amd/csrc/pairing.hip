// Curvature-signature sort, pruned sample-pair search and vote emission.
//
// Pairing restates src/cpu/pairing.cpp:14-35: all selected-ordinal pairs
// a < b with |dkmin| <= t*M, |dkmax| <= t*M (M = max of the four |k| and
// 1e-10) and |p_a - p_b| >= 0.03*box, in (a asc, b asc) order.
//   1. gather  : selected ids -> ordinal-order sample SoA (14 lanes) and
//                order-preserving uint64 keys of kmin (-0 canonicalised);
//   2. sort    : LDS-staged stable LSD radix sort of (key, ordinal);
//   3. windows : per source a conservative kmin interval (a superset of every
//                b the predicate can accept) found by binary search;
//   4. count   : one wavefront per source walks its window with coalesced
//                loads of (ordinal, kmin, kmax) and counts accepted b > a;
//   5. scan    : exclusive prefix of the counts -> output offsets; the total
//                is read back once and checked against the pair budget
//                before any output is written;
//   6. fill    : the same walk appends accepted ordinals to an LDS list;
//                each element's rank in the list (count of smaller ordinals)
//                is its position, which restores the reference's order;
//                sources with more than kCap pairs walk the ordinals in order
//                instead (k_pair_long), so no list can overflow.
// Vote emission restates src/cpu/transformations.cpp:17-73 per pair: an
// optional mirror vote (two dot tests > .75) then two rigid votes
// R = F_b(+-) F_a^T, t = q - R p.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>

#include "common.hpp"
#include "session.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;  // sources per block
constexpr int kCap = 512;                 // LDS order-restoration capacity per source
constexpr double kEps = 2.220446049250313e-16;

// Sample SoA lanes (ordinal order).
enum { L_X = 0, L_Y, L_Z, L_K0, L_K1, L_NX, L_NY, L_NZ, L_DX, L_DY, L_DZ, L_XX, L_XY, L_XZ, L_N };

__device__ inline uint64_t order_key(double v) {
  if (v == 0.0)
    v = 0.0;  // -0 and +0 compare equal: give them one key
  uint64_t u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}

// Flags: 1 = selected id out of range, 2 = non-finite coordinate/curvature.
__global__ void k_gather_samples(const double *x, const double *y, const double *z,
                                 const double *feat, uint64_t n, const int32_t *sel, uint64_t S,
                                 double cx, double cy, double cz, double *s64, float *s32,
                                 uint64_t *keys, uint32_t *vals, unsigned *flags) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= S)
    return;
  int32_t id = sel[i];
  double v[L_N];
  if (id < 0 || static_cast<uint64_t>(id) >= n) {
    atomicOr(flags, 1u);
    for (int k = 0; k < L_N; ++k)
      v[k] = 0;
  } else {
    v[L_X] = x[id];
    v[L_Y] = y[id];
    v[L_Z] = z[id];
    v[L_K0] = feat[9 * n + id];
    v[L_K1] = feat[10 * n + id];
    for (int k = 0; k < 9; ++k)
      v[L_NX + k] = feat[k * n + id];
    bool fin = isfinite(v[L_X]) && isfinite(v[L_Y]) && isfinite(v[L_Z]) && isfinite(v[L_K0]) &&
               isfinite(v[L_K1]);
    if (!fin)
      atomicOr(flags, 2u);
  }
  for (int k = 0; k < L_N; ++k)
    s64[k * S + i] = v[k];
  if (s32) {
    // FP32 copy: positions relative to the model centre
    s32[L_X * S + i] = static_cast<float>(v[L_X] - cx);
    s32[L_Y * S + i] = static_cast<float>(v[L_Y] - cy);
    s32[L_Z * S + i] = static_cast<float>(v[L_Z] - cz);
    for (int k = L_K0; k < L_N; ++k)
      s32[k * S + i] = static_cast<float>(v[k]);
  }
  keys[i] = order_key(v[L_K0]);
  vals[i] = static_cast<uint32_t>(i);
}

__global__ void k_sorted_gather(const uint32_t *ord, const double *s64, const float *s32,
                                uint64_t S, double *sk, float *sk32) {
  uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= S)
    return;
  uint32_t o = ord[j];
  sk[j] = s64[L_K0 * S + o];
  sk[S + j] = s64[L_K1 * S + o];
  if (sk32) {
    sk32[j] = s32[L_K0 * S + o];
    sk32[S + j] = s32[L_K1 * S + o];
  }
}

// Conservative kmin window.  If a pair passes, with Ma = max(|ka|,|Ka|,1e-10)
// and M the pair magnitude: M <= Ma/(1-t) (if the other sample dominates,
// its magnitude m satisfies m <= Ma + t m), hence |ka - kb| <= t Ma/(1-t).
// t is inflated by 16 eps to cover the reference's rounding of the
// difference and of t*M, the radius by another 16 eps, and the endpoints by
// 16 eps of their magnitude; t >= .5 or denormal-range t use the full range.
__global__ void k_windows(const double *s64, const double *sk, uint64_t S, double t,
                          uint64_t *win) {
  uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= S)
    return;
  uint64_t lo = 0, hi = S;
  double tp = t * (1 + 16 * kEps);
  if (tp < 0.5 && t > 1e-290) {
    double a = s64[L_K0 * S + i], b = s64[L_K1 * S + i];
    double ma = fmax(fmax(fabs(a), fabs(b)), 1e-10);
    double r = tp * ma / (1 - tp) * (1 + 16 * kEps);
    double pad = 16 * kEps * (fabs(a) + r);
    double low = a - r - pad, high = a + r + pad;
    // lower_bound(low): first j with !(sk[j] < low)
    uint64_t f = 0, c = S;
    while (c) {
      uint64_t h = c / 2, m = f + h;
      if (sk[m] < low) {
        f = m + 1;
        c -= h + 1;
      } else
        c = h;
    }
    lo = f;
    // upper_bound(high): first j with sk[j] > high
    f = 0;
    c = S;
    while (c) {
      uint64_t h = c / 2, m = f + h;
      if (!(sk[m] > high)) {
        f = m + 1;
        c -= h + 1;
      } else
        c = h;
    }
    hi = f;
  }
  win[i] = lo;
  win[S + i] = hi;
}

// Exact reference predicate in FP64 (pairing.cpp:22-28 as restated above).
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
// gk bounds the relative error of the curvature tests, gd of the distance.
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
  const double *s64;
  const float *s32;
  const double *sk;   // sorted kmin | kmax (2*S)
  const float *sk32;  // sorted FP32 kmin | kmax
  const uint32_t *ord;
  const uint64_t *win;
  uint64_t S;
  double t, dist;
  float t32, dist32, gk, gd;
  unsigned long long *stats;  // [0] candidates, [1] FP64 rechecks
};

// Decide candidate (source i, sorted slot j holding ordinal o) in the chosen arithmetic.
template <bool kGuarded>
__device__ inline bool decide(const PairArgs &A, uint64_t i, uint64_t j, uint32_t o,
                              double a0, double a1, float f0, float f1, unsigned &rechecks) {
  const uint64_t S = A.S;
  if (!kGuarded) {
    double b0 = A.sk[j], b1 = A.sk[S + j];
    double m = fmax(fmax(fabs(a0), fabs(a1)), fmax(fmax(fabs(b0), fabs(b1)), 1e-10));
    double tol = A.t * m;
    if (fabs(a0 - b0) > tol || fabs(a1 - b1) > tol)
      return false;
    return accept64(a0, a1, b0, b1, A.s64[L_X * S + i], A.s64[L_Y * S + i], A.s64[L_Z * S + i],
                    A.s64[L_X * S + o], A.s64[L_Y * S + o], A.s64[L_Z * S + o], A.t, A.dist);
  } else {
    int r = screen32(f0, f1, A.sk32[j], A.sk32[S + j], A.t32, A.gk);
    if (r == 0)
      return false;
    if (r == 1) {
      int d = screen32_dist(A.s32[L_X * S + i] - A.s32[L_X * S + o],
                            A.s32[L_Y * S + i] - A.s32[L_Y * S + o],
                            A.s32[L_Z * S + i] - A.s32[L_Z * S + o], A.dist32, A.gd);
      if (d != 2)
        return d == 1;
    }
    ++rechecks;
    return accept64(a0, a1, A.sk[j], A.sk[S + j], A.s64[L_X * S + i], A.s64[L_Y * S + i],
                    A.s64[L_Z * S + i], A.s64[L_X * S + o], A.s64[L_Y * S + o],
                    A.s64[L_Z * S + o], A.t, A.dist);
  }
}

template <bool kGuarded>
__global__ void __launch_bounds__(kThreads)
    k_pair_count(PairArgs A, uint64_t *counts, uint32_t *long_list, unsigned *long_count) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / kWave;
  if (i >= A.S)
    return;
  const uint64_t S = A.S;
  const uint64_t lo = A.win[i], hi = A.win[S + i];
  const double a0 = A.s64[L_K0 * S + i], a1 = A.s64[L_K1 * S + i];
  const float f0 = kGuarded ? A.s32[L_K0 * S + i] : 0.f, f1 = kGuarded ? A.s32[L_K1 * S + i] : 0.f;
  uint64_t cnt = 0;
  unsigned rechecks = 0;
  for (uint64_t j0 = lo; j0 < hi; j0 += kWave) {
    uint64_t j = j0 + lane;
    bool ok = false;
    if (j < hi) {
      uint32_t o = A.ord[j];
      if (o > i)
        ok = decide<kGuarded>(A, i, j, o, a0, a1, f0, f1, rechecks);
    }
    cnt += __popcll(ballot(ok));
  }
  if (lane == 0) {
    counts[i] = cnt;
    if (cnt > static_cast<uint64_t>(kCap)) {
      unsigned slot = atomicAdd(long_count, 1u);
      long_list[slot] = static_cast<uint32_t>(i);
    }
    atomicAdd(&A.stats[0], static_cast<unsigned long long>(hi - lo));
  }
  rechecks = wave_sum(rechecks);
  if (lane == 0 && rechecks)
    atomicAdd(&A.stats[1], static_cast<unsigned long long>(rechecks));
}

template <bool kGuarded>
__global__ void __launch_bounds__(kThreads)
    k_pair_fill(PairArgs A, const uint64_t *counts, const uint64_t *offsets, const int32_t *sel,
                int32_t *out_a, int32_t *out_b) {
  __shared__ uint32_t list[kWaves][kCap];
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kWaves + wave;
  if (i >= A.S)
    return;
  const uint64_t S = A.S;
  const uint64_t total = counts[i];
  if (total == 0 || total > static_cast<uint64_t>(kCap))
    return;
  const uint64_t lo = A.win[i], hi = A.win[S + i];
  const double a0 = A.s64[L_K0 * S + i], a1 = A.s64[L_K1 * S + i];
  const float f0 = kGuarded ? A.s32[L_K0 * S + i] : 0.f, f1 = kGuarded ? A.s32[L_K1 * S + i] : 0.f;
  uint32_t n = 0;
  unsigned rechecks = 0;
  uint32_t *L = list[wave];
  for (uint64_t j0 = lo; j0 < hi; j0 += kWave) {
    uint64_t j = j0 + lane;
    bool ok = false;
    uint32_t o = 0;
    if (j < hi) {
      o = A.ord[j];
      if (o > i)
        ok = decide<kGuarded>(A, i, j, o, a0, a1, f0, f1, rechecks);
    }
    uint64_t mask = ballot(ok);
    uint32_t pos = n + static_cast<uint32_t>(lanes_below(mask));
    if (ok && pos < kCap)
      L[pos] = o;
    n += static_cast<uint32_t>(__popcll(mask));
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  // count and fill make identical decisions; the clamps only keep every
  // access inside this source's output range should they ever disagree
  if (n > kCap)
    n = kCap;
  if (n > total)
    n = static_cast<uint32_t>(total);
  const uint64_t base = offsets[i];
  const int32_t src = sel ? sel[i] : static_cast<int32_t>(i);
  for (uint32_t e = lane; e < n; e += kWave) {
    uint32_t v = L[e], r = 0;
    for (uint32_t k = 0; k < n; ++k)
      r += L[k] < v;
    out_a[base + r] = src;
    out_b[base + r] = sel ? sel[v] : static_cast<int32_t>(v);
  }
}

// Sources with more than kCap pairs: walk the targets in ordinal order.
template <bool kGuarded>
__global__ void __launch_bounds__(kThreads)
    k_pair_long(PairArgs A, const uint32_t *long_list, unsigned long_count,
                const uint64_t *offsets, const int32_t *sel, int32_t *out_a, int32_t *out_b) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kWaves + threadIdx.x / kWave;
  if (w >= long_count)
    return;
  const uint64_t S = A.S;
  const uint64_t i = long_list[w];
  const double a0 = A.s64[L_K0 * S + i], a1 = A.s64[L_K1 * S + i];
  const double ax = A.s64[L_X * S + i], ay = A.s64[L_Y * S + i], az = A.s64[L_Z * S + i];
  const int32_t src = sel ? sel[i] : static_cast<int32_t>(i);
  uint64_t cursor = offsets[i];
  const uint64_t end = offsets[i + 1];
  for (uint64_t b0 = i + 1; b0 < S && cursor < end; b0 += kWave) {
    uint64_t b = b0 + lane;
    bool ok = false;
    if (b < S)
      ok = accept64(a0, a1, A.s64[L_K0 * S + b], A.s64[L_K1 * S + b], ax, ay, az,
                    A.s64[L_X * S + b], A.s64[L_Y * S + b], A.s64[L_Z * S + b], A.t, A.dist);
    uint64_t mask = ballot(ok);
    uint64_t p = cursor + lanes_below(mask);
    if (ok && p < end) {
      out_a[p] = src;
      out_b[p] = sel ? sel[b] : static_cast<int32_t>(b);
    }
    cursor += __popcll(mask);
  }
}

// ------------------------------------------------------------ vote geometry
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

// Rotation matrix -> unit quaternion (w, x, y, z), w >= 0 (Shepperd's method).
template <class R> __device__ inline void quat(const R *m, R *q) {
  R tr = m[0] + m[4] + m[8];
  R w, x, y, z;
  if (tr > R(0)) {
    R s = sqrt(tr + R(1)) * R(2);
    w = R(0.25) * s;
    x = (m[7] - m[5]) / s;
    y = (m[2] - m[6]) / s;
    z = (m[3] - m[1]) / s;
  } else if (m[0] > m[4] && m[0] > m[8]) {
    R s = sqrt(R(1) + m[0] - m[4] - m[8]) * R(2);
    w = (m[7] - m[5]) / s;
    x = R(0.25) * s;
    y = (m[1] + m[3]) / s;
    z = (m[2] + m[6]) / s;
  } else if (m[4] > m[8]) {
    R s = sqrt(R(1) + m[4] - m[0] - m[8]) * R(2);
    w = (m[2] - m[6]) / s;
    x = (m[1] + m[3]) / s;
    y = R(0.25) * s;
    z = (m[5] + m[7]) / s;
  } else {
    R s = sqrt(R(1) + m[8] - m[0] - m[4]) * R(2);
    w = (m[3] - m[1]) / s;
    x = (m[2] + m[6]) / s;
    y = (m[5] + m[7]) / s;
    z = R(0.25) * s;
  }
  if (w < R(0))
    w = -w, x = -x, y = -y, z = -z;
  q[0] = w, q[1] = x, q[2] = y, q[3] = z;
}

// FP32 vector helpers for the guarded emission.
struct F3 {
  float x, y, z;
};
__device__ inline F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float fcomp(F3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Loads of one sample's 14 lanes (ordinal order).
__device__ inline V3 lane3(const double *s, uint64_t S, int l, uint64_t i) {
  return v3(s[l * S + i], s[(l + 1) * S + i], s[(l + 2) * S + i]);
}
__device__ inline F3 lane3f(const float *s, uint64_t S, int l, uint64_t i) {
  return f3(s[l * S + i], s[(l + 1) * S + i], s[(l + 2) * S + i]);
}

struct VoteArgs {
  const double *s64;
  const float *s32;
  uint64_t S;
  const int32_t *a, *b;  // pair ordinals
  uint64_t P;
  double cx, cy, cz;
  float grefl, gdom;       // guard bands (FP32 path)
  unsigned long long *stats;  // [2] reflection rechecks, [3] axis-tie rechecks
};

// Reflection decision and plane for pair k in the chosen arithmetic.  In the
// guarded path the FP32 result is used unless a test value lies inside its
// guard band, in which case the FP64 decision is taken.
template <bool kGuarded>
__device__ inline bool pair_reflection(const VoteArgs &A, uint64_t k, unsigned &rr, unsigned &rd) {
  const uint64_t S = A.S;
  uint64_t i = static_cast<uint64_t>(A.a[k]), o = static_cast<uint64_t>(A.b[k]);
  if (!kGuarded) {
    Geo64 g = geo64(lane3(A.s64, S, L_X, i), lane3(A.s64, S, L_X, o), lane3(A.s64, S, L_NX, i),
                    lane3(A.s64, S, L_DX, i), lane3(A.s64, S, L_XX, i), lane3(A.s64, S, L_NX, o),
                    lane3(A.s64, S, L_DX, o), lane3(A.s64, S, L_XX, o));
    return g.refl;
  }
  F3 p = lane3f(A.s32, S, L_X, i), q = lane3f(A.s32, S, L_X, o);
  F3 dv = q - p;
  float len = sqrtf(fdot(dv, dv));
  bool tie = false;
  F3 n = len > 1e-14f ? dv * (1.0f / len) : f3(0, 0, 0);
  {
    float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
    float m1 = fmaxf(ax, fmaxf(ay, az));
    float m2 = fmaxf(fminf(ax, ay), fminf(fmaxf(ax, ay), az));
    tie = m1 - m2 <= A.gdom || !(len > 1e-6f);
  }
  if (!tie) {
    int dom = 0;
    for (int kk = 1; kk < 3; ++kk)
      if (fabsf(fcomp(n, kk)) > fabsf(fcomp(n, dom)))
        dom = kk;
    if (fcomp(n, dom) < 0)
      n = n * -1.0f;
    F3 an = lane3f(A.s32, S, L_NX, i), ad = lane3f(A.s32, S, L_DX, i);
    F3 bn = lane3f(A.s32, S, L_NX, o), bd = lane3f(A.s32, S, L_DX, o);
    // (I - 2 n n^T) v = v - 2 (n.v) n
    F3 ma = an - n * (2.0f * fdot(n, an)), md = ad - n * (2.0f * fdot(n, ad));
    float r1 = fdot(ma, bn), r2 = fabsf(fdot(md, bd));
    if (fabsf(r1 - 0.75f) > A.grefl && fabsf(r2 - 0.75f) > A.grefl)
      return r1 > 0.75f && r2 > 0.75f;
    if (r1 < 0.75f - A.grefl || r2 < 0.75f - A.grefl)
      return false;  // one test fails outside its band
    ++rr;
  } else {
    ++rd;
  }
  Geo64 g = geo64(lane3(A.s64, S, L_X, i), lane3(A.s64, S, L_X, o), lane3(A.s64, S, L_NX, i),
                  lane3(A.s64, S, L_DX, i), lane3(A.s64, S, L_XX, i), lane3(A.s64, S, L_NX, o),
                  lane3(A.s64, S, L_DX, o), lane3(A.s64, S, L_XX, o));
  return g.refl;
}

constexpr int kVoteItems = 4;
constexpr int kVoteTile = kThreads * kVoteItems;  // pairs per block

template <bool kGuarded>
__global__ void __launch_bounds__(kThreads) k_vote_count(VoteArgs A, uint32_t *block_refl) {
  __shared__ uint32_t wsum[kWaves];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kVoteTile;
  unsigned rr = 0, rd = 0;
  uint32_t c = 0;
  for (int r = 0; r < kVoteItems; ++r) {
    uint64_t k = base + r * kThreads + threadIdx.x;
    if (k < A.P)
      c += pair_reflection<kGuarded>(A, k, rr, rd);
  }
  c = wave_sum(c);
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == 0)
    wsum[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    block_refl[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

template <class R> struct VoteOut {
  R *q4;        // rigid: 4 lanes, stride 2P
  R *t3;        // rigid: 3 lanes, stride 2P
  int32_t *rp;  // rigid pair ids
  R *n3;        // reflection: 3 lanes, stride Rtot
  R *d1;        // reflection offsets
  int32_t *fp;  // reflection pair ids
  uint64_t R_total;
};

template <bool kGuarded, class R>
__global__ void __launch_bounds__(kThreads)
    k_vote_emit(VoteArgs A, const uint64_t *block_off, VoteOut<R> out) {
  __shared__ uint32_t wsum[kWaves + 1];
  const uint64_t S = A.S, P = A.P, P2 = 2 * A.P;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kVoteTile;
  uint64_t run = block_off[blockIdx.x];
  unsigned rr = 0, rd = 0;
  const double cx = A.cx, cy = A.cy, cz = A.cz;
  for (int r = 0; r < kVoteItems; ++r) {
    uint64_t k = base + r * kThreads + threadIdx.x;
    bool live = k < P;
    bool refl = live && pair_reflection<kGuarded>(A, k, rr, rd);
    // block-wide exclusive prefix of refl in pair order
    uint64_t m = ballot(refl);
    uint32_t below = static_cast<uint32_t>(lanes_below(m));
    if (lane == 0)
      wsum[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w)
      before += wsum[w];
    uint32_t tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    uint64_t slot = run + before + below;
    run += tot;
    if (!live)
      continue;
    uint64_t i = static_cast<uint64_t>(A.a[k]), o = static_cast<uint64_t>(A.b[k]);
    R rq[2][4], rt[2][3];
    R pn[3], pd = 0;
    if (!kGuarded) {
      Geo64 g = geo64(lane3(A.s64, S, L_X, i), lane3(A.s64, S, L_X, o), lane3(A.s64, S, L_NX, i),
                      lane3(A.s64, S, L_DX, i), lane3(A.s64, S, L_XX, i), lane3(A.s64, S, L_NX, o),
                      lane3(A.s64, S, L_DX, o), lane3(A.s64, S, L_XX, o));
      V3 c = v3(cx, cy, cz);
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
      pn[0] = static_cast<R>(g.n.x), pn[1] = static_cast<R>(g.n.y), pn[2] = static_cast<R>(g.n.z);
      pd = static_cast<R>(g.d + dot(g.n, c));
    } else {
      // FP32 frames in centre-relative coordinates: t_c = q' - R p'
      F3 p = lane3f(A.s32, S, L_X, i), q = lane3f(A.s32, S, L_X, o);
      F3 an = lane3f(A.s32, S, L_NX, i), ad = lane3f(A.s32, S, L_DX, i),
         ax = lane3f(A.s32, S, L_XX, i);
      F3 bn = lane3f(A.s32, S, L_NX, o), bd = lane3f(A.s32, S, L_DX, o),
         bx = lane3f(A.s32, S, L_XX, o);
      for (int sg = 0; sg < 2; ++sg) {
        float s = sg == 0 ? 1.0f : -1.0f;
        F3 b0 = bd * s, b1 = bx * s;
        // R = [b0 b1 bn] [ad ax an]^T
        float m[9];
        float Bc[3][3] = {{b0.x, b1.x, bn.x}, {b0.y, b1.y, bn.y}, {b0.z, b1.z, bn.z}};
        float Ac[3][3] = {{ad.x, ax.x, an.x}, {ad.y, ax.y, an.y}, {ad.z, ax.z, an.z}};
        for (int rI = 0; rI < 3; ++rI)
          for (int cI = 0; cI < 3; ++cI)
            m[rI * 3 + cI] = Bc[rI][0] * Ac[cI][0] + Bc[rI][1] * Ac[cI][1] + Bc[rI][2] * Ac[cI][2];
        float rp0 = m[0] * p.x + m[1] * p.y + m[2] * p.z;
        float rp1 = m[3] * p.x + m[4] * p.y + m[5] * p.z;
        float rp2 = m[6] * p.x + m[7] * p.y + m[8] * p.z;
        float qq[4];
        quat<float>(m, qq);
        for (int kk = 0; kk < 4; ++kk)
          rq[sg][kk] = static_cast<R>(qq[kk]);
        rt[sg][0] = static_cast<R>(q.x - rp0), rt[sg][1] = static_cast<R>(q.y - rp1),
        rt[sg][2] = static_cast<R>(q.z - rp2);
      }
      if (refl) {
        F3 dv = q - p;
        float len = sqrtf(fdot(dv, dv));
        F3 n = dv * (1.0f / len);
        float ax_ = fabsf(n.x), ay_ = fabsf(n.y), az_ = fabsf(n.z);
        float m1 = fmaxf(ax_, fmaxf(ay_, az_));
        float m2 = fmaxf(fminf(ax_, ay_), fminf(fmaxf(ax_, ay_), az_));
        if (m1 - m2 <= A.gdom || !(len > 1e-6f)) {
          Geo64 g = geo64(lane3(A.s64, S, L_X, i), lane3(A.s64, S, L_X, o),
                          lane3(A.s64, S, L_NX, i), lane3(A.s64, S, L_DX, i),
                          lane3(A.s64, S, L_XX, i), lane3(A.s64, S, L_NX, o),
                          lane3(A.s64, S, L_DX, o), lane3(A.s64, S, L_XX, o));
          V3 c = v3(cx, cy, cz);
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
          // plane offset about the centre: -n . (p' + q')/2
          pn[0] = static_cast<R>(n.x), pn[1] = static_cast<R>(n.y), pn[2] = static_cast<R>(n.z);
          pd = static_cast<R>(-fdot(n, (p + q) * 0.5f));
        }
      } else {
        pn[0] = pn[1] = pn[2] = R(0);
      }
    }
    // rigid slots 2k, 2k+1 (sign +1 then -1)
    for (int kk = 0; kk < 4; ++kk) {
      out.q4[kk * P2 + 2 * k] = rq[0][kk];
      out.q4[kk * P2 + 2 * k + 1] = rq[1][kk];
    }
    for (int kk = 0; kk < 3; ++kk) {
      out.t3[kk * P2 + 2 * k] = rt[0][kk];
      out.t3[kk * P2 + 2 * k + 1] = rt[1][kk];
    }
    out.rp[2 * k] = static_cast<int32_t>(k);
    out.rp[2 * k + 1] = static_cast<int32_t>(k);
    if (refl && slot < out.R_total) {
      const uint64_t RT = out.R_total;
      out.n3[slot] = pn[0];
      out.n3[RT + slot] = pn[1];
      out.n3[2 * RT + slot] = pn[2];
      out.d1[slot] = pd;
      out.fp[slot] = static_cast<int32_t>(k);
    }
  }
  rr = wave_sum(rr);
  rd = wave_sum(rd);
  if (lane == 0 && (rr || rd)) {
    atomicAdd(&A.stats[2], static_cast<unsigned long long>(rr));
    atomicAdd(&A.stats[3], static_cast<unsigned long long>(rd));
  }
}

// --------------------------------------- reference-layout transformations
// Backend::transformations output: 28 lanes in global vote order.
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

__global__ void k_read_total(const uint64_t *off, uint64_t n, uint64_t *dst) { *dst = off[n]; }

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

// Read back a few device words (one synchronisation).
static void read_back(Session *s, void *dst, const void *src, size_t bytes) {
  void *h = host_stage(s, bytes);
  MVS_HIP(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, s->stream));
  MVS_HIP(hipStreamSynchronize(s->stream));
  std::memcpy(dst, h, bytes);
}

// Guard-band widths of the FP32 screen (relative).  Each bounds the rounding
// of the FP32 evaluation (inputs rounded once to FP32, a handful of FP32
// operations) with a wide margin; see the derivations next to the kernels.
static float guard_curvature(double t) {
  return static_cast<float>(1e-4 + 16.0 * 5.960464477539063e-08 * (1.0 + 1.0 / t));
}
static float guard_distance(double box, double dist) {
  return static_cast<float>(1e-4 + 16.0 * 5.960464477539063e-08 * box / dist);
}

struct PairState {
  double *s64;
  float *s32;
  double *sk;
  float *sk32;
  uint32_t *ord;
  uint64_t *win, *counts, *offsets;
  uint32_t *long_list;
  unsigned long long *stats;
};

static PairState pair_state(Session *s, uint64_t S, bool guarded) {
  PairState st{};
  st.s64 = scratch_as<double>(s, kSlotSamples, L_N * S);
  st.s32 = guarded ? scratch_as<float>(s, kSlotPlanSamples32, L_N * S) : nullptr;
  st.sk = scratch_as<double>(s, kSlotPlanSortedK, 2 * S);
  st.sk32 = guarded ? scratch_as<float>(s, kSlotPlanSortedK32, 2 * S) : nullptr;
  st.ord = scratch_as<uint32_t>(s, kSlotPlanOrd, S);
  st.win = scratch_as<uint64_t>(s, kSlotPlanWin, 2 * S);
  st.counts = scratch_as<uint64_t>(s, kSlotPlanCounts, S + 1);
  st.offsets = scratch_as<uint64_t>(s, kSlotPlanOffsets, S + 1);
  st.long_list = scratch_as<uint32_t>(s, kSlotPlanLong, S + 1);
  st.stats = scratch_as<unsigned long long>(s, kSlotPlanStats, 8);
  return st;
}

static PairArgs pair_args(const PairState &st, uint64_t S, double t, double box, bool guarded) {
  PairArgs A{};
  A.s64 = st.s64;
  A.s32 = st.s32;
  A.sk = st.sk;
  A.sk32 = st.sk32;
  A.ord = st.ord;
  A.win = st.win;
  A.S = S;
  A.t = t;
  A.dist = 0.03 * box;
  A.t32 = static_cast<float>(t);
  A.dist32 = static_cast<float>(A.dist);
  A.gk = guarded ? guard_curvature(t) : 0.f;
  A.gd = guarded ? guard_distance(box, A.dist) : 0.f;
  A.stats = st.stats;
  return A;
}

__global__ void k_zero_words(unsigned long long *p, int n) {
  if (threadIdx.x < n)
    p[threadIdx.x] = 0;
}

// count phase shared by the host and device layers
void pair_count_device(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                       const double *feat, uint64_t S, const int32_t *sel, double box, double t,
                       uint64_t max_pairs, int precision, const double *center, uint64_t *P_out) {
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
  if (S < 2) {
    plan.P = 0;
    plan.valid = true;
    *P_out = 0;
    // keep the sample SoA for vote emission of an empty pair set
    return;
  }
  const bool guarded = precision == MVS_FP32_GUARDED;
  PairState st = pair_state(s, S, guarded);
  uint64_t *keys = scratch_as<uint64_t>(s, kSlotPlanKeys, 2 * S);
  uint32_t *vals = scratch_as<uint32_t>(s, kSlotSort3, S);
  unsigned *flags = scratch_as<unsigned>(s, kSlotFlags, 4);
  MVS_HIP(hipMemsetAsync(flags, 0, 4 * sizeof(unsigned), s->stream));
  MVS_LAUNCH(s, "zero_stats", k_zero_words, 1, 64, 0, st.stats, 8);
  double cx = center ? center[0] : 0, cy = center ? center[1] : 0, cz = center ? center[2] : 0;
  MVS_LAUNCH(s, "pair_gather", k_gather_samples, blocks_for(S, kThreads), kThreads, 0, x, y, z,
             feat, n, sel, S, cx, cy, cz, st.s64, st.s32, keys, vals, flags);
  radix_sort_pairs_u64(s, keys, vals, keys + S, st.ord, S, 64);
  MVS_LAUNCH(s, "pair_sorted_gather", k_sorted_gather, blocks_for(S, kThreads), kThreads, 0,
             st.ord, st.s64, st.s32, S, st.sk, st.sk32);
  MVS_LAUNCH(s, "pair_windows", k_windows, blocks_for(S, kThreads), kThreads, 0, st.s64, st.sk, S,
             t, st.win);
  unsigned *long_count = flags + 1;
  PairArgs A = pair_args(st, S, t, box, guarded);
  unsigned grid = blocks_for(S, kWaves);
  if (guarded)
    MVS_LAUNCH(s, "pair_count", k_pair_count<true>, grid, kThreads, 0, A, st.counts, st.long_list,
               long_count);
  else
    MVS_LAUNCH(s, "pair_count", k_pair_count<false>, grid, kThreads, 0, A, st.counts,
               st.long_list, long_count);
  exclusive_scan_u64(s, st.counts, st.offsets, S);
  // read back flags, long count and the total in one synchronisation
  uint64_t *tot = reinterpret_cast<uint64_t *>(flags + 2);
  MVS_LAUNCH(s, "read_total", k_read_total, 1, 1, 0, st.offsets, S, tot);
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
  PairArgs A = pair_args(st, S, plan.threshold, plan.box, guarded);
  const int32_t *sel = ordinals ? nullptr : plan.selected;
  unsigned grid = blocks_for(S, kWaves);
  if (guarded)
    MVS_LAUNCH(s, "pair_fill", k_pair_fill<true>, grid, kThreads, 0, A, st.counts, st.offsets, sel,
               a, b);
  else
    MVS_LAUNCH(s, "pair_fill", k_pair_fill<false>, grid, kThreads, 0, A, st.counts, st.offsets,
               sel, a, b);
  if (plan.long_sources) {
    unsigned lg = blocks_for(plan.long_sources, kWaves);
    if (guarded)
      MVS_LAUNCH(s, "pair_long", k_pair_long<true>, lg, kThreads, 0, A, st.long_list,
                 static_cast<unsigned>(plan.long_sources), st.offsets, sel, a, b);
    else
      MVS_LAUNCH(s, "pair_long", k_pair_long<false>, lg, kThreads, 0, A, st.long_list,
                 static_cast<unsigned>(plan.long_sources), st.offsets, sel, a, b);
  }
}

static VoteArgs vote_args(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                          const double *center, bool guarded) {
  PairPlan &plan = s->pair_plan;
  if (!plan.valid)
    raise(MVS_EINVAL, "vote emission needs the sample plan of a pair count");
  if (guarded && plan.precision != MVS_FP32_GUARDED)
    raise(MVS_EINVAL, "FP32-guarded votes need an FP32-guarded pair plan");
  const uint64_t S = plan.S;
  VoteArgs A{};
  A.s64 = scratch_as<double>(s, kSlotSamples, L_N * (S ? S : 1));
  A.s32 = guarded ? scratch_as<float>(s, kSlotPlanSamples32, L_N * (S ? S : 1)) : nullptr;
  A.S = S;
  A.a = a;
  A.b = b;
  A.P = P;
  A.cx = center ? center[0] : 0;
  A.cy = center ? center[1] : 0;
  A.cz = center ? center[2] : 0;
  // |q' - p'| >= 0.03 box for every pair; FP32 centre-relative coordinates
  // are bounded by box, so the direction carries a relative error of at
  // most ~ 8 eps32 box / (0.03 box); the bands are 20x wider than that.
  A.grefl = 2e-4f;
  A.gdom = 2e-4f;
  A.stats = scratch_as<unsigned long long>(s, kSlotPlanStats, 8);
  return A;
}

void vote_count_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b, int precision,
                       uint64_t *R_out) {
  const bool guarded = precision == MVS_FP32_GUARDED;
  VoteArgs A = vote_args(s, P, a, b, nullptr, guarded);
  unsigned nb = blocks_for(P ? P : 1, kVoteTile);
  uint32_t *blk = scratch_as<uint32_t>(s, kSlotVoteBlock, nb);
  uint64_t *off = scratch_as<uint64_t>(s, kSlotVoteBlockOff, nb + 1);
  if (P) {
    if (guarded)
      MVS_LAUNCH(s, "vote_count", k_vote_count<true>, nb, kThreads, 0, A, blk);
    else
      MVS_LAUNCH(s, "vote_count", k_vote_count<false>, nb, kThreads, 0, A, blk);
  }
  exclusive_scan_u32(s, blk, off, P ? nb : 0);
  uint64_t *tot = scratch_as<uint64_t>(s, kSlotFlags, 4) + 1;
  MVS_LAUNCH(s, "read_total", k_read_total, 1, 1, 0, off, static_cast<uint64_t>(P ? nb : 0), tot);
  uint64_t R;
  read_back(s, &R, tot, sizeof R);
  s->vote_plan.P = P;
  s->vote_plan.R = R;
  s->vote_plan.precision = precision;
  s->vote_plan.valid = true;
  *R_out = R;
}

void vote_emit_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                      const double *center, int precision, void *q4, void *t3, int32_t *rp,
                      void *n3, void *d1, int32_t *fp) {
  const bool guarded = precision == MVS_FP32_GUARDED;
  VotePlan &vp = s->vote_plan;
  if (!vp.valid || vp.P != P || vp.precision != precision)
    raise(MVS_EINVAL, "vote emission without a matching vote count");
  if (P == 0)
    return;
  VoteArgs A = vote_args(s, P, a, b, center, guarded);
  unsigned nb = blocks_for(P, kVoteTile);
  uint64_t *off = scratch_as<uint64_t>(s, kSlotVoteBlockOff, nb + 1);
  if (guarded) {
    VoteOut<float> o{static_cast<float *>(q4), static_cast<float *>(t3), rp,
                     static_cast<float *>(n3), static_cast<float *>(d1), fp, vp.R};
    MVS_LAUNCH(s, "vote_emit", (k_vote_emit<true, float>), nb, kThreads, 0, A, off, o);
  } else {
    VoteOut<double> o{static_cast<double *>(q4), static_cast<double *>(t3), rp,
                      static_cast<double *>(n3), static_cast<double *>(d1), fp, vp.R};
    MVS_LAUNCH(s, "vote_emit", (k_vote_emit<false, double>), nb, kThreads, 0, A, off, o);
  }
}

void pair_stats(Session *s, unsigned long long out[4]) {
  read_back(s, out, scratch_as<unsigned long long>(s, kSlotPlanStats, 8), 4 * sizeof(uint64_t));
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_pair_count(mvs_session *sp, uint64_t n, const double *x, const double *y,
                       const double *z, const double *feat, uint64_t S, const int32_t *sel,
                       double box, double t, uint64_t max_pairs, int precision,
                       uint64_t *pairs) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!pairs)
      raise(MVS_EINVAL, "null output pointer");
    pair_count_device(s, n, x, y, z, feat, S, sel, box, t, max_pairs, precision, nullptr, pairs);
  });
}

int mvs_dev_pair_fill(mvs_session *sp, int32_t *a, int32_t *b, int ordinals) {
  return guarded([&] { pair_fill_device(as_session(sp), a, b, ordinals != 0); });
}

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
    vote_emit_device(as_session(sp), P, a, b, center, precision, q4, t3, rp, n3, d1, fp);
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
    MVS_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned), s->stream));
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
    uint64_t P = 0;
    pair_count_device(s, n, dv.ptr, dv.ptr + n, dv.ptr + 2 * n, feat.ptr, S, dsel.ptr, box, t,
                      max_pairs, precision, nullptr, &P);
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
