/*
 * CPU oracle (test infrastructure only; see oracle.h for the pinning list).
 * Plain C11, FP64, single thread.  Built with -ffp-contract=off so that every
 * product and sum rounds as written, like the reference CPU build.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char g_error[256];
const char *oracle_last_error(void) { return g_error; }
static int fail(int code, const char *msg) {
  snprintf(g_error, sizeof g_error, "%s", msg);
  return code;
}

/* ---------------------------------------------------------------- math ---
 * src/core/math.cpp:9-21, 23-38, 113-119 */
typedef struct {
  double x, y, z;
} vec3;
static vec3 vadd(vec3 a, vec3 b) { return (vec3){a.x + b.x, a.y + b.y, a.z + b.z}; }
static vec3 vsub(vec3 a, vec3 b) { return (vec3){a.x - b.x, a.y - b.y, a.z - b.z}; }
static vec3 vscale(vec3 a, double s) { return (vec3){a.x * s, a.y * s, a.z * s}; }
/* math.cpp:12: division by a scalar is a reciprocal multiply. */
static vec3 vdiv(vec3 a, double s) { return vscale(a, 1 / s); }
static double vdot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static vec3 vcross(vec3 a, vec3 b) {
  return (vec3){a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static double vnorm(vec3 a) { return sqrt(vdot(a, a)); }
static vec3 vunit(vec3 a) {
  double n = vnorm(a);
  return n > 1e-14 ? vdiv(a, n) : (vec3){0, 0, 0};
}
static double vget(vec3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
/* row-major 3x3 */
typedef struct {
  double m[9];
} mat3;
static vec3 mapply(const mat3 *m, vec3 v) {
  return (vec3){m->m[0] * v.x + m->m[1] * v.y + m->m[2] * v.z,
                m->m[3] * v.x + m->m[4] * v.y + m->m[5] * v.z,
                m->m[6] * v.x + m->m[7] * v.y + m->m[8] * v.z};
}
static mat3 mtranspose(const mat3 *a) {
  mat3 r = {{a->m[0], a->m[3], a->m[6], a->m[1], a->m[4], a->m[7], a->m[2],
             a->m[5], a->m[8]}};
  return r;
}
/* math.cpp:31-38: r starts at zero and accumulates k = 0, 1, 2. */
static mat3 mmul(const mat3 *a, const mat3 *b) {
  mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k)
        s += a->m[i * 3 + k] * b->m[k * 3 + j];
      r.m[i * 3 + j] = s;
    }
  return r;
}
/* math.cpp:113-119 */
static mat3 mreflection(vec3 n) {
  mat3 r = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i * 3 + j] -= 2 * vget(n, i) * vget(n, j);
  return r;
}
/* transformations.cpp:14-16 / clustering.cpp:21-23: a, b, c are columns. */
static mat3 mframe(vec3 a, vec3 b, vec3 c) {
  mat3 r = {{a.x, b.x, c.x, a.y, b.y, c.y, a.z, b.z, c.z}};
  return r;
}

/* ------------------------------------------------------------- mt19937_64 ---
 * The engine the C++ standard specifies ([rand.predef]); state layout and
 * refill schedule follow libstdc++ (x[312], index p, refill when p >= 312). */
#define MT_N 312
#define MT_M 156
void or_mt64_seed(or_mt64 *g, uint64_t seed) {
  g->x[0] = seed;
  for (uint64_t i = 1; i < MT_N; ++i)
    g->x[i] = 6364136223846793005ULL * (g->x[i - 1] ^ (g->x[i - 1] >> 62)) + i;
  g->p = MT_N;
}
static void mt_refill(or_mt64 *g) {
  const uint64_t upper = ~0ULL << 31, lower = ~upper, a = 0xB5026F5AA96619E9ULL;
  uint64_t *x = g->x;
  for (int k = 0; k < MT_N; ++k) {
    uint64_t y = (x[k] & upper) | (x[(k + 1) % MT_N] & lower);
    x[k] = x[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1) ? a : 0);
  }
  g->p = 0;
}
uint64_t or_mt64_next(or_mt64 *g) {
  if (g->p >= MT_N)
    mt_refill(g);
  uint64_t z = g->x[g->p++];
  z ^= (z >> 29) & 0x5555555555555555ULL;
  z ^= (z << 17) & 0x71D67FFFEDA60000ULL;
  z ^= (z << 37) & 0xFFF7EEE000000000ULL;
  z ^= z >> 43;
  return z;
}

/* ------------------------------------------------------------- prepare ---
 * src/core/prepare.cpp:13-58 (topology) and :102-149 (prepare). */
typedef struct {
  int32_t lo, hi, face, a, b, order;
} edge_entry;
static int cmp_edge_entry(const void *pa, const void *pb) {
  const edge_entry *a = pa, *b = pb;
  if (a->lo != b->lo)
    return a->lo < b->lo ? -1 : 1;
  if (a->hi != b->hi)
    return a->hi < b->hi ? -1 : 1;
  return a->order < b->order ? -1 : (a->order > b->order);
}
static int cmp_i32(const void *pa, const void *pb) {
  int32_t a = *(const int32_t *)pa, b = *(const int32_t *)pb;
  return a < b ? -1 : (a > b);
}

int oracle_prepare(int64_t nv, const double *x, const double *y, const double *z,
                   int64_t nf, const int32_t *f0, const int32_t *f1, const int32_t *f2,
                   or_prepared *o) {
  if (nv <= 0 || nf <= 0)
    return fail(OR_ERUNTIME, "Mesh is empty");
  if (nv > 2147483647LL || nf > 2147483647LL)
    return fail(OR_ERUNTIME, "Mesh exceeds 32-bit index capacity");
  vec3 lo = {x[0], y[0], z[0]}, hi = lo;
  for (int64_t i = 0; i < nv; ++i) {
    double v[3] = {x[i], y[i], z[i]};
    double *l = &lo.x, *h = &hi.x;
    for (int k = 0; k < 3; ++k) {
      if (!isfinite(v[k]))
        return fail(OR_ERUNTIME, "Non-finite coordinate");
      /* std::min / std::max */
      l[k] = v[k] < l[k] ? v[k] : l[k];
      h[k] = h[k] < v[k] ? v[k] : h[k];
    }
  }
  o->box_size = vnorm(vsub(hi, lo));
  if (o->box_size < 1e-14)
    return fail(OR_ERUNTIME, "Mesh has zero extent");
  o->rejected_faces = o->nonmanifold_edges = o->orientation_conflicts = 0;
  int64_t F = 0;
  for (int64_t i = 0; i < nf; ++i) {
    int32_t f[3] = {f0[i], f1[i], f2[i]};
    for (int k = 0; k < 3; ++k)
      if (f[k] < 0 || f[k] >= nv)
        return fail(OR_ERUNTIME, "Face index out of range");
    vec3 a = {x[f[0]], y[f[0]], z[f[0]]}, b = {x[f[1]], y[f[1]], z[f[1]]},
         c = {x[f[2]], y[f[2]], z[f[2]]};
    if (vnorm(vcross(vsub(b, a), vsub(c, a))) <= 1e-14 * o->box_size * o->box_size) {
      ++o->rejected_faces;
      continue;
    }
    o->f0[F] = f[0];
    o->f1[F] = f[1];
    o->f2[F] = f[2];
    ++F;
  }
  o->face_count = F;
  if (!F)
    return fail(OR_ERUNTIME, "No nondegenerate triangles");

  /* topology(): normals and weights accumulate in face order, corner order. */
  for (int64_t i = 0; i < nv; ++i) {
    o->nx[i] = o->ny[i] = o->nz[i] = 0;
    o->weights[i] = 0;
  }
  int32_t *ring_count = calloc((size_t)nv, sizeof(int32_t));
  int32_t *inc_count = calloc((size_t)nv, sizeof(int32_t));
  edge_entry *ent = malloc((size_t)(3 * F) * sizeof(edge_entry));
  for (int64_t i = 0; i < F; ++i) {
    int32_t f[3] = {o->f0[i], o->f1[i], o->f2[i]};
    vec3 p0 = {x[f[0]], y[f[0]], z[f[0]]}, p1 = {x[f[1]], y[f[1]], z[f[1]]},
         p2 = {x[f[2]], y[f[2]], z[f[2]]};
    vec3 n = vcross(vsub(p1, p0), vsub(p2, p0));
    double area = vnorm(n) / 2;
    n = vunit(n);
    o->fnx[i] = n.x;
    o->fny[i] = n.y;
    o->fnz[i] = n.z;
    for (int k = 0; k < 3; ++k) {
      int32_t a = f[k], b = f[(k + 1) % 3];
      vec3 acc = {o->nx[a], o->ny[a], o->nz[a]};
      acc = vadd(acc, n);
      o->nx[a] = acc.x;
      o->ny[a] = acc.y;
      o->nz[a] = acc.z;
      o->weights[a] += area / 3;
      ring_count[a]++;
      ring_count[b]++;
      inc_count[a]++;
      edge_entry e;
      e.lo = a < b ? a : b;
      e.hi = a < b ? b : a;
      e.face = (int32_t)i;
      e.a = a;
      e.b = b;
      e.order = (int32_t)(3 * i + k);
      ent[3 * i + k] = e;
    }
  }
  /* ring rows (with duplicates) then sort + unique per row */
  uint64_t *roff = malloc((size_t)(nv + 1) * sizeof(uint64_t));
  roff[0] = 0;
  for (int64_t i = 0; i < nv; ++i)
    roff[i + 1] = roff[i] + (uint64_t)ring_count[i];
  int32_t *rbuf = malloc((size_t)roff[nv] * sizeof(int32_t) + 1);
  int32_t *rfill = calloc((size_t)nv, sizeof(int32_t));
  o->inc_off[0] = 0;
  for (int64_t i = 0; i < nv; ++i)
    o->inc_off[i + 1] = o->inc_off[i] + (uint64_t)inc_count[i];
  int32_t *ifill = calloc((size_t)nv, sizeof(int32_t));
  for (int64_t i = 0; i < F; ++i) {
    int32_t f[3] = {o->f0[i], o->f1[i], o->f2[i]};
    for (int k = 0; k < 3; ++k) {
      int32_t a = f[k], b = f[(k + 1) % 3];
      rbuf[roff[a] + (uint64_t)rfill[a]++] = b;
      rbuf[roff[b] + (uint64_t)rfill[b]++] = a;
      o->inc_idx[o->inc_off[a] + (uint64_t)ifill[a]++] = (int32_t)i;
    }
  }
  o->inc_nnz = (int64_t)o->inc_off[nv];
  int64_t nnz = 0;
  o->nbr_off[0] = 0;
  for (int64_t i = 0; i < nv; ++i) {
    vec3 acc = {o->nx[i], o->ny[i], o->nz[i]};
    acc = vunit(acc);
    o->nx[i] = acc.x;
    o->ny[i] = acc.y;
    o->nz[i] = acc.z;
    int32_t *row = rbuf + roff[i];
    int32_t len = ring_count[i];
    qsort(row, (size_t)len, sizeof(int32_t), cmp_i32);
    for (int32_t k = 0; k < len; ++k)
      if (k == 0 || row[k] != row[k - 1])
        o->nbr_idx[nnz++] = row[k];
    o->nbr_off[i + 1] = (uint64_t)nnz;
  }
  o->nbr_nnz = nnz;
  /* edges in (lo, hi) order; entries of one edge keep face/corner order */
  qsort(ent, (size_t)(3 * F), sizeof(edge_entry), cmp_edge_entry);
  int64_t E = 0;
  for (int64_t s = 0; s < 3 * F;) {
    int64_t t = s + 1;
    while (t < 3 * F && ent[t].lo == ent[s].lo && ent[t].hi == ent[s].hi)
      ++t;
    int64_t size = t - s;
    if (size > 2)
      ++o->nonmanifold_edges;
    if (size == 2 && ent[s].a != ent[s + 1].b)
      ++o->orientation_conflicts;
    or_edge e = {ent[s].lo, ent[s].hi, ent[s].a, ent[s].b, ent[s].face,
                 size == 2 ? ent[s + 1].face : -1};
    o->edges[E++] = e;
    s = t;
  }
  o->edge_count = E;
  free(ring_count);
  free(inc_count);
  free(ent);
  free(roff);
  free(rbuf);
  free(rfill);
  free(ifill);
  if (o->nonmanifold_edges || o->orientation_conflicts)
    return fail(OR_ERUNTIME, "Mesh must have manifold edges and consistent "
                             "face winding; repair input mesh first");
  /* prepare.cpp:144-147: center = (sum of vertices) / N, reciprocal multiply */
  vec3 c = {0, 0, 0};
  for (int64_t i = 0; i < nv; ++i)
    c = vadd(c, (vec3){x[i], y[i], z[i]});
  c = vdiv(c, (double)nv);
  o->center[0] = c.x;
  o->center[1] = c.y;
  o->center[2] = c.z;
  return OR_OK;
}

/* ----------------------------------------------------------- curvature ---
 * src/cpu/curvature.cpp:14-79 */
int oracle_curvature(int64_t nv, const double *x, const double *y, const double *z,
                     const double *nx, const double *ny, const double *nz,
                     const double *fnx, const double *fny, const double *fnz,
                     const uint64_t *nbr_off, const int32_t *nbr_idx,
                     int64_t ne, const or_edge *edges, int smoothing,
                     double *feat, uint8_t *valid) {
  for (int64_t k = 0; k < 11 * nv; ++k)
    feat[k] = 0;
  for (int64_t i = 0; i < nv; ++i)
    valid[i] = 0;
  if (nv <= 0)
    return OR_OK;
  double *T = calloc((size_t)(9 * nv), sizeof(double));
  double *N = malloc((size_t)(9 * nv) * sizeof(double));
  int *count = calloc((size_t)nv, sizeof(int));
  double mean = 0;
  for (int64_t e = 0; e < ne; ++e) {
    vec3 h = {x[edges[e].hi], y[edges[e].hi], z[edges[e].hi]};
    vec3 l = {x[edges[e].lo], y[edges[e].lo], z[edges[e].lo]};
    mean += vnorm(vsub(h, l));
  }
  mean /= (double)ne;
  for (int64_t e = 0; e < ne; ++e) {
    const or_edge *E = &edges[e];
    if (E->face1 < 0)
      continue;
    vec3 d = vsub((vec3){x[E->b], y[E->b], z[E->b]}, (vec3){x[E->a], y[E->a], z[E->a]});
    double len = vnorm(d);
    d = vunit(d);
    vec3 n1 = {fnx[E->face0], fny[E->face0], fnz[E->face0]};
    vec3 n2 = {fnx[E->face1], fny[E->face1], fnz[E->face1]};
    double c = vdot(n1, n2);
    c = c < -1.0 ? -1.0 : (1.0 < c ? 1.0 : c); /* std::clamp */
    double beta = atan2(vdot(vcross(n1, n2), d), c);
    int32_t ends[2] = {E->a, E->b};
    for (int q = 0; q < 2; ++q) {
      int32_t v = ends[q];
      ++count[v];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          T[9 * v + i * 3 + j] += vget(d, i) * vget(d, j) * beta * len / mean;
    }
  }
  for (int64_t i = 0; i < nv; ++i) {
    int c = count[i] > 1 ? count[i] : 1;
    for (int k = 0; k < 9; ++k)
      T[9 * i + k] /= c;
  }
  for (int it = 0; it < smoothing; ++it) {
    for (int64_t i = 0; i < nv; ++i) {
      for (int k = 0; k < 9; ++k)
        N[9 * i + k] = T[9 * i + k];
      for (uint64_t j = nbr_off[i]; j < nbr_off[i + 1]; ++j)
        for (int k = 0; k < 9; ++k)
          N[9 * i + k] += T[9 * (int64_t)nbr_idx[j] + k];
      double deg1 = (double)(nbr_off[i + 1] - nbr_off[i] + 1);
      for (int k = 0; k < 9; ++k)
        N[9 * i + k] /= deg1;
    }
    double *s = T;
    T = N;
    N = s;
  }
  for (int64_t i = 0; i < nv; ++i) {
    vec3 n = {nx[i], ny[i], nz[i]};
    if (vnorm(n) < 0.5 || nbr_off[i + 1] - nbr_off[i] < 3)
      continue;
    vec3 axis = fabs(n.x) < 0.8 ? (vec3){1, 0, 0} : (vec3){0, 1, 0};
    vec3 u = vunit(vcross(n, axis)), v = vcross(n, u);
    mat3 t;
    memcpy(t.m, T + 9 * i, sizeof t.m);
    double a = vdot(u, mapply(&t, u)), b = vdot(u, mapply(&t, v)),
           c = vdot(v, mapply(&t, v));
    double angle = 0.5 * atan2(2 * b, a - c);
    vec3 high = vadd(vscale(u, cos(angle)), vscale(v, sin(angle)));
    vec3 low = vunit(vcross(high, n));
    double delta = hypot(a - c, 2 * b);
    double kmin = (a + c - delta) / 2, kmax = (a + c + delta) / 2;
    vec3 dmax = vunit(vcross(n, low));
    double vals[11] = {n.x, n.y, n.z, low.x, low.y, low.z, dmax.x, dmax.y, dmax.z, kmin, kmax};
    for (int k = 0; k < 11; ++k)
      feat[k * nv + i] = vals[k];
    valid[i] = isfinite(kmin) && isfinite(kmax) && delta > 1e-10;
  }
  free(T);
  free(N);
  free(count);
  return OR_OK;
}

/* ------------------------------------------------------------ sampling ---
 * src/core/sampling.cpp:5-24: exponential keys, sorted by (key, index). */
typedef struct {
  double key;
  int32_t idx;
} key_entry;
static int cmp_key(const void *pa, const void *pb) {
  const key_entry *a = pa, *b = pb;
  if (a->key < b->key)
    return -1;
  if (b->key < a->key)
    return 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}
int oracle_sample_vertices(int64_t nv, const double *w, double ratio, or_mt64 *rng,
                           int32_t *out, int64_t *count) {
  key_entry *keys = malloc((size_t)(nv ? nv : 1) * sizeof(key_entry));
  int64_t nk = 0;
  for (int64_t i = 0; i < nv; ++i)
    if (w[i] > 0) {
      double u = ((double)(or_mt64_next(rng) >> 11) + 0.5) / 9007199254740992.0;
      keys[nk].key = -log(u) / w[i];
      keys[nk].idx = (int32_t)i;
      ++nk;
    }
  qsort(keys, (size_t)nk, sizeof(key_entry), cmp_key);
  int64_t target = (int64_t)ceil((double)nv * ratio);
  if (target > nk)
    target = nk;
  for (int64_t i = 0; i < target; ++i)
    out[i] = keys[i].idx;
  qsort(out, (size_t)target, sizeof(int32_t), cmp_i32);
  *count = target;
  free(keys);
  return OR_OK;
}

/* --------------------------------------------------------------- prune ---
 * src/cpu/selection.cpp:21-27 */
int oracle_prune(int64_t nv, const double *kmin, const double *kmax,
                 const uint8_t *valid, int64_t ns, const int32_t *sample,
                 double prune, int32_t *out, int64_t *count) {
  int64_t c = 0;
  for (int64_t s = 0; s < ns; ++s) {
    int32_t i = sample[s];
    if (i < 0 || i >= nv)
      return fail(OR_EINVAL, "sample vertex index out of range");
    double a = fabs(kmin[i]), b = fabs(kmax[i]);
    double maximum = a < b ? b : a;
    if (valid[i] && maximum > 1e-10 && fabs(kmin[i] - kmax[i]) / maximum > 1 - prune)
      out[c++] = i;
  }
  *count = c;
  return OR_OK;
}

/* ------------------------------------------------------------- pairing ---
 * src/cpu/pairing.cpp:14-35 (brute force over ordinals a < b). */
/* Sources [a0, a1) of the pair loop; the pairs of source a follow those of
 * source a - 1, so consecutive ranges concatenate into the CPU order. */
static int pairing_range(int64_t nv, const double *x, const double *y, const double *z,
                         const double *kmin, const double *kmax, int64_t ns,
                         const int32_t *sel, double box, double t, uint64_t max_pairs,
                         int64_t cap, int32_t *out_a, int32_t *out_b, int64_t *count,
                         int64_t a0, int64_t a1) {
  int64_t c = 0;
  double dist = 0.03 * box;
  for (int64_t a = a0; a < a1; ++a) {
    int32_t i = sel[a];
    double pk0 = kmin[i], pk1 = kmax[i];
    vec3 pi = {x[i], y[i], z[i]};
    for (int64_t b = a + 1; b < ns; ++b) {
      int32_t j = sel[b];
      double qk0 = kmin[j], qk1 = kmax[j];
      /* std::max over an initializer list: first maximal element */
      double m = fabs(pk0);
      double cand[4] = {fabs(pk1), fabs(qk0), fabs(qk1), 1e-10};
      for (int k = 0; k < 4; ++k)
        if (m < cand[k])
          m = cand[k];
      double tol = t * m;
      if (fabs(pk0 - qk0) > tol || fabs(pk1 - qk1) > tol ||
          vnorm(vsub(pi, (vec3){x[j], y[j], z[j]})) < dist)
        continue;
      if ((uint64_t)c >= max_pairs)
        return fail(OR_EBUDGET, "Pair budget exceeded; lower --sample-ratio "
                                "or tighten --pair-threshold");
      if (c >= cap)
        return fail(OR_EOVERFLOW, "pair output capacity exceeded");
      out_a[c] = i;
      out_b[c] = j;
      ++c;
    }
  }
  *count = c;
  return OR_OK;
}
int oracle_pairing(int64_t nv, const double *x, const double *y, const double *z,
                   const double *kmin, const double *kmax, int64_t ns,
                   const int32_t *sel, double box, double t, uint64_t max_pairs,
                   int64_t cap, int32_t *out_a, int32_t *out_b, int64_t *count) {
  for (int64_t s = 0; s < ns; ++s)
    if (sel[s] < 0 || sel[s] >= nv)
      return fail(OR_EINVAL, "selected vertex index out of range");
  return pairing_range(nv, x, y, z, kmin, kmax, ns, sel, box, t, max_pairs, cap, out_a, out_b,
                       count, 0, ns);
}

/* ----------------------------------------------------- transformations ---
 * src/cpu/transformations.cpp:17-73 */
static void put_vote(double *lanes, int64_t cap, int64_t row, const mat3 *r, vec3 t,
                     vec3 n, double d, const double *desc) {
  for (int k = 0; k < 9; ++k)
    lanes[k * cap + row] = r->m[k];
  double v[7] = {t.x, t.y, t.z, n.x, n.y, n.z, d};
  for (int k = 0; k < 7; ++k)
    lanes[(9 + k) * cap + row] = v[k];
  for (int k = 0; k < 12; ++k)
    lanes[(16 + k) * cap + row] = desc[k];
}
int oracle_transformations(int64_t nv, const double *x, const double *y,
                           const double *z, const double *feat, int64_t np,
                           const int32_t *pa, const int32_t *pb, double box,
                           const double *center, uint64_t max_votes, int64_t cap,
                           int32_t *pair, uint8_t *kind, double *lanes,
                           int64_t *count) {
  vec3 ctr = {center[0], center[1], center[2]};
  int64_t c = 0;
#define FEAT(lane, i) feat[(int64_t)(lane)*nv + (i)]
#define FV(lane, i) ((vec3){FEAT(lane, i), FEAT(lane + 1, i), FEAT(lane + 2, i)})
  for (int64_t i = 0; i < np; ++i) {
    if (pa[i] < 0 || pb[i] < 0 || pa[i] >= nv || pb[i] >= nv)
      return fail(OR_EINVAL, "pair vertex index out of range");
    int32_t ia = pa[i], ib = pb[i];
    vec3 p = {x[ia], y[ia], z[ia]}, q = {x[ib], y[ib], z[ib]};
    vec3 an = FV(0, ia), ad = FV(3, ia), ax = FV(6, ia);
    vec3 bn = FV(0, ib), bd = FV(3, ib), bx = FV(6, ib);
    vec3 n = vunit(vsub(q, p));
    int dom = 0;
    for (int k = 1; k < 3; ++k)
      if (fabs(vget(n, k)) > fabs(vget(n, dom)))
        dom = k;
    if (vget(n, dom) < 0)
      n = vscale(n, -1);
    double d = -vdot(n, vdiv(vadd(p, q), 2));
    mat3 mirror = mreflection(n);
    if (vdot(mapply(&mirror, an), bn) > 0.75 &&
        fabs(vdot(mapply(&mirror, ad), bd)) > 0.75) {
      if ((uint64_t)c >= max_votes)
        return fail(OR_EBUDGET, "Vote budget exceeded; lower --sample-ratio or "
                                "tighten --pair-threshold");
      if (c >= cap)
        return fail(OR_EOVERFLOW, "vote output capacity exceeded");
      double desc[12] = {n.x, n.y, n.z, (d + vdot(n, ctr)) / box};
      put_vote(lanes, cap, c, &mirror, vscale(n, -2 * d), n, d, desc);
      pair[c] = (int32_t)i;
      kind[c] = 1;
      ++c;
    }
    mat3 fa = mframe(ad, ax, an), fat = mtranspose(&fa);
    for (int s = 0; s < 2; ++s) {
      double sign = s == 0 ? 1.0 : -1.0;
      mat3 fb = mframe(vscale(bd, sign), vscale(bx, sign), bn);
      mat3 r = mmul(&fb, &fat);
      vec3 t = vsub(q, mapply(&r, p));
      vec3 cen = vdiv(vsub(vadd(mapply(&r, ctr), t), ctr), box);
      double desc[12];
      for (int k = 0; k < 9; ++k)
        desc[k] = r.m[k] / 2;
      desc[9] = cen.x;
      desc[10] = cen.y;
      desc[11] = cen.z;
      if ((uint64_t)c >= max_votes)
        return fail(OR_EBUDGET, "Vote budget exceeded; lower --sample-ratio or "
                                "tighten --pair-threshold");
      if (c >= cap)
        return fail(OR_EOVERFLOW, "vote output capacity exceeded");
      put_vote(lanes, cap, c, &r, t, (vec3){0, 0, 0}, 0, desc);
      pair[c] = (int32_t)i;
      kind[c] = 0;
      ++c;
    }
  }
#undef FV
#undef FEAT
  *count = c;
  return OR_OK;
}

/* ---------------------------------------------------------- clustering ---
 * src/cpu/clustering.cpp:24-208 */
typedef struct {
  int32_t vote;
  uint64_t count;
} counter_entry;
typedef struct {
  counter_entry *e;
  int64_t n, cap;
} counter;
typedef struct {
  double v[12];
  int dim;
  int reflection;
} center_t;

static double ddist(const double *a, const double *b, int dim) {
  double s = 0;
  for (int i = 0; i < dim; ++i) {
    double d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

typedef struct {
  int64_t nv;
  const uint8_t *kind;
  const double *desc;
  int32_t *touched;  /* votes in the current trajectory's support */
  int64_t ntouched;
  uint64_t *support; /* dense per-vote iteration counts, zero between calls */
} ms_work;

/* clustering.cpp:156-185 (trajectory) with the evaluator of :186-208. */
static void trajectory(ms_work *w, double *mean, int dim, int refl, double h2,
                       int max_iterations, int *converged, int *exhausted) {
  *converged = *exhausted = 0;
  double sum[12];
  for (int it = 0; it < max_iterations; ++it) {
    double weight = 0;
    for (int k = 0; k < dim; ++k)
      sum[k] = 0;
    for (int64_t j = 0; j < w->nv; ++j) {
      if ((w->kind[j] != 0) != refl)
        continue;
      double s = 0;
      for (int k = 0; k < dim; ++k) {
        double d = mean[k] - w->desc[(int64_t)k * w->nv + j];
        s += d * d;
      }
      if (s < h2) {
        double wt = 1 - s / h2;
        weight += wt;
        if (w->support[j]++ == 0)
          w->touched[w->ntouched++] = (int32_t)j;
        for (int k = 0; k < dim; ++k)
          sum[k] += wt * w->desc[(int64_t)k * w->nv + j];
      }
    }
    if (weight <= 1e-14) {
      *exhausted = 1;
      break;
    }
    for (int k = 0; k < dim; ++k)
      sum[k] /= weight;
    double movement = ddist(mean, sum, dim);
    for (int k = 0; k < dim; ++k)
      mean[k] = sum[k];
    if (movement < 1e-6 * h2) {
      *converged = 1;
      break;
    }
  }
}

typedef struct {
  int reflection;
  double rot[9], trans[3], pn[3], po;
  int32_t *votes;
  int64_t nvotes;
  int64_t order;
} cluster_out;
static int cmp_cluster(const void *pa, const void *pb) {
  const cluster_out *a = pa, *b = pb;
  if (a->nvotes != b->nvotes)
    return a->nvotes > b->nvotes ? -1 : 1; /* descending size */
  return a->order < b->order ? -1 : (a->order > b->order); /* stable */
}

int oracle_clustering(int64_t nv, const uint8_t *kind, const double *desc,
                      const double *rotation, double box, const double *center,
                      double bandwidth, int max_iterations, int max_clusters,
                      or_mt64 *rng, int32_t *n_clusters, uint8_t *c_refl,
                      double *c_rot, double *c_trans, double *c_pn, double *c_po,
                      uint64_t *member_off, int32_t *members, int32_t *limited) {
  double h = bandwidth, h2 = h * h;
  int64_t nc = 0, ccap = 16;
  center_t *centers = malloc((size_t)ccap * sizeof(center_t));
  counter *counters = malloc((size_t)ccap * sizeof(counter));
  uint8_t *visited = calloc((size_t)(nv ? nv : 1), 1);
  int32_t *seeds = malloc((size_t)(nv ? nv : 1) * sizeof(int32_t));
  ms_work w = {nv, kind, desc, malloc((size_t)(nv ? nv : 1) * sizeof(int32_t)), 0,
               calloc((size_t)(nv ? nv : 1), sizeof(uint64_t))};
  uint64_t support_entries = 0;
  int32_t lim = 0;
  int rc = OR_OK;
  for (;;) {
    int64_t ns = 0;
    for (int64_t i = 0; i < nv; ++i)
      if (!visited[i])
        seeds[ns++] = (int32_t)i;
    if (!ns)
      break;
    int32_t seed = seeds[or_mt64_next(rng) % (uint64_t)ns];
    int refl = kind[seed] != 0;
    int dim = refl ? 4 : 12;
    double mean[12] = {0};
    for (int k = 0; k < 12; ++k)
      mean[k] = desc[(int64_t)k * nv + seed];
    int conv, exh;
    w.ntouched = 0;
    trajectory(&w, mean, dim, refl, h2, max_iterations, &conv, &exh);
    /* support as (vote, count) sorted by vote (std::map order) */
    qsort(w.touched, (size_t)w.ntouched, sizeof(int32_t), cmp_i32);
    counter sup = {malloc((size_t)(w.ntouched ? w.ntouched : 1) * sizeof(counter_entry)),
                   w.ntouched, w.ntouched};
    for (int64_t k = 0; k < w.ntouched; ++k) {
      int32_t j = w.touched[k];
      sup.e[k].vote = j;
      sup.e[k].count = w.support[j];
      w.support[j] = 0;
      visited[j] = 1;
    }
    if (exh)
      visited[seed] = 1;
    if (!conv)
      ++lim;
    int64_t merge = -1;
    for (int64_t c = 0; c < nc; ++c)
      if (centers[c].reflection == refl && ddist(mean, centers[c].v, dim) < h2 / 4) {
        merge = c;
        break;
      }
    if (merge < 0) {
      if (nc == ccap) {
        ccap *= 2;
        centers = realloc(centers, (size_t)ccap * sizeof(center_t));
        counters = realloc(counters, (size_t)ccap * sizeof(counter));
      }
      memset(&centers[nc], 0, sizeof(center_t));
      for (int k = 0; k < dim; ++k)
        centers[nc].v[k] = mean[k];
      centers[nc].dim = dim;
      centers[nc].reflection = refl;
      support_entries += (uint64_t)sup.n;
      counters[nc++] = sup;
    } else {
      double old = (double)counters[merge].n, now = (double)sup.n;
      if (old + now > 0)
        for (int k = 0; k < dim; ++k)
          centers[merge].v[k] = (old * centers[merge].v[k] + now * mean[k]) / (old + now);
      counter *m = &counters[merge];
      int64_t before = m->n;
      counter_entry *merged =
          malloc((size_t)(m->n + sup.n ? m->n + sup.n : 1) * sizeof(counter_entry));
      int64_t i = 0, j = 0, k = 0;
      while (i < m->n || j < sup.n) {
        if (j >= sup.n || (i < m->n && m->e[i].vote < sup.e[j].vote))
          merged[k++] = m->e[i++];
        else if (i >= m->n || sup.e[j].vote < m->e[i].vote)
          merged[k++] = sup.e[j++];
        else {
          merged[k] = m->e[i++];
          merged[k++].count += sup.e[j++].count;
        }
      }
      free(m->e);
      m->e = merged;
      m->n = k;
      m->cap = k;
      support_entries += (uint64_t)(m->n - before);
      free(sup.e);
    }
    if (support_entries > 5000000) {
      rc = fail(OR_ERUNTIME, "Mean-shift support budget exceeded; lower --sample-ratio");
      goto done;
    }
    if (nc > 20000) {
      rc = fail(OR_ERUNTIME, "Too many mean-shift modes; increase "
                             "--bandwidth or lower --sample-ratio");
      goto done;
    }
  }
  {
    /* winners: strictly larger count wins; earlier cluster keeps ties */
    uint64_t *best = calloc((size_t)(nv ? nv : 1), sizeof(uint64_t));
    int32_t *winner = malloc((size_t)(nv ? nv : 1) * sizeof(int32_t));
    for (int64_t i = 0; i < nv; ++i)
      winner[i] = -1;
    for (int64_t c = 0; c < nc; ++c)
      for (int64_t k = 0; k < counters[c].n; ++k) {
        counter_entry e = counters[c].e[k];
        if (e.count > best[e.vote]) {
          best[e.vote] = e.count;
          winner[e.vote] = (int32_t)c;
        }
      }
    int64_t *msize = calloc((size_t)(nc ? nc : 1), sizeof(int64_t));
    for (int64_t i = 0; i < nv; ++i)
      if (winner[i] >= 0)
        msize[winner[i]]++;
    cluster_out *outs = malloc((size_t)(nc ? nc : 1) * sizeof(cluster_out));
    int64_t no = 0;
    vec3 origin = {center[0], center[1], center[2]};
    for (int64_t c = 0; c < nc; ++c) {
      if (!msize[c])
        continue;
      cluster_out *o = &outs[no];
      memset(o, 0, sizeof *o);
      o->order = no;
      o->reflection = centers[c].reflection;
      o->votes = malloc((size_t)msize[c] * sizeof(int32_t));
      for (int64_t i = 0; i < nv; ++i)
        if (winner[i] == c)
          o->votes[o->nvotes++] = (int32_t)i;
      const double *dsc = centers[c].v;
      if (o->reflection) {
        vec3 n = {dsc[0], dsc[1], dsc[2]};
        double length = vnorm(n);
        vec3 pn = vunit(n);
        double po = dsc[3] * box / length - vdot(pn, origin);
        mat3 r = mreflection(pn);
        vec3 t = vscale(pn, -2 * po);
        memcpy(o->rot, r.m, sizeof o->rot);
        o->trans[0] = t.x;
        o->trans[1] = t.y;
        o->trans[2] = t.z;
        o->pn[0] = pn.x;
        o->pn[1] = pn.y;
        o->pn[2] = pn.z;
        o->po = po;
      } else {
        mat3 m;
        for (int k = 0; k < 9; ++k)
          m.m[k] = dsc[k] * 2;
        vec3 c0 = {m.m[0], m.m[3], m.m[6]}, c1 = {m.m[1], m.m[4], m.m[7]};
        vec3 a = vunit(c0);
        vec3 b = vunit(vsub(c1, vscale(a, vdot(a, c1))));
        if (vnorm(a) < 0.5 || vnorm(b) < 0.5) {
          int64_t v0 = o->votes[0];
          for (int k = 0; k < 9; ++k)
            m.m[k] = rotation[(int64_t)k * nv + v0];
        } else
          m = mframe(a, b, vcross(a, b));
        vec3 t = {dsc[9], dsc[10], dsc[11]};
        vec3 tr = vsub(vadd(vscale(t, box), origin), mapply(&m, origin));
        memcpy(o->rot, m.m, sizeof o->rot);
        o->trans[0] = tr.x;
        o->trans[1] = tr.y;
        o->trans[2] = tr.z;
      }
      ++no;
    }
    qsort(outs, (size_t)no, sizeof(cluster_out), cmp_cluster);
    int64_t keep = no < max_clusters ? no : max_clusters;
    member_off[0] = 0;
    for (int64_t c = 0; c < keep; ++c) {
      cluster_out *o = &outs[c];
      c_refl[c] = (uint8_t)o->reflection;
      for (int k = 0; k < 9; ++k)
        c_rot[9 * c + k] = o->rot[k];
      for (int k = 0; k < 3; ++k) {
        c_trans[3 * c + k] = o->trans[k];
        c_pn[3 * c + k] = o->pn[k];
      }
      c_po[c] = o->po;
      for (int64_t k = 0; k < o->nvotes; ++k)
        members[member_off[c] + (uint64_t)k] = o->votes[k];
      member_off[c + 1] = member_off[c] + (uint64_t)o->nvotes;
    }
    *n_clusters = (int32_t)keep;
    *limited = lim;
    for (int64_t c = 0; c < no; ++c)
      free(outs[c].votes);
    free(outs);
    free(msize);
    free(best);
    free(winner);
  }
done:
  for (int64_t c = 0; c < nc; ++c)
    free(counters[c].e);
  free(counters);
  free(centers);
  free(visited);
  free(seeds);
  free(w.touched);
  free(w.support);
  return rc;
}

/* ------------------------------------------------ batched-seed mean-shift ---
 * No reference counterpart (the reference runs sequential seeds only).  Each
 * seed follows m <- m + sum w (x - m) / sum w with the shifted, truncated
 * Gaussian weight w = exp(-d^2 / 2h^2) - exp(-c^2 / 2) for d < c h, which is
 * the mean-shift ascent of the density with profile derivative -w. */
int oracle_meanshift_batched(int64_t nv, int dim, const double *pts,
                             const uint8_t *kind, int dim_rigid, int dim_refl,
                             int64_t ns, const double *seeds,
                             const uint8_t *seed_kind, double h, double cutoff,
                             int iterations, double *out, double *weight) {
  if (dim < 1 || dim > 12 || dim_rigid > dim || dim_refl > dim)
    return fail(OR_EINVAL, "bad dimension");
  double r2 = cutoff * h * cutoff * h, inv = 1 / (2 * h * h);
  double floor_w = exp(-0.5 * cutoff * cutoff);
  for (int64_t s = 0; s < ns; ++s) {
    double m[12] = {0}, acc[12];
    int refl = seed_kind[s] != 0, d = refl ? dim_refl : dim_rigid;
    for (int k = 0; k < dim; ++k)
      m[k] = seeds[(int64_t)k * ns + s];
    double wsum = 0;
    for (int it = 0; it < iterations; ++it) {
      wsum = 0;
      for (int k = 0; k < d; ++k)
        acc[k] = 0;
      for (int64_t j = 0; j < nv; ++j) {
        if ((kind[j] != 0) != refl)
          continue;
        double d2 = 0;
        for (int k = 0; k < d; ++k) {
          double e = pts[(int64_t)k * nv + j] - m[k];
          d2 += e * e;
        }
        if (d2 < r2) {
          double wj = exp(-d2 * inv) - floor_w;
          wsum += wj;
          for (int k = 0; k < d; ++k)
            acc[k] += wj * (pts[(int64_t)k * nv + j] - m[k]);
        }
      }
      if (wsum > 0)
        for (int k = 0; k < d; ++k)
          m[k] += acc[k] / wsum;
    }
    for (int k = 0; k < dim; ++k)
      out[(int64_t)k * ns + s] = m[k];
    weight[s] = wsum;
  }
  return OR_OK;
}

/* ------------------------------------------------------- verification ---
 * src/core/math.cpp:65-112: squared distance from p to segment ab and to
 * triangle abc (Voronoi regions of Ericson, Real-Time Collision Detection,
 * 5.1.5; degenerate triangles fall back to their three edges). */
static double or_clamp01(double v) { return v < 0.0 ? 0.0 : (1.0 < v ? 1.0 : v); }
static double or_min(double a, double b) { return b < a ? b : a; }
static double or_seg_d2(vec3 p, vec3 a, vec3 b) {
  vec3 d = vsub(b, a);
  double l = vdot(d, d);
  double t = l > 1e-14 ? or_clamp01(vdot(vsub(p, a), d) / l) : 0;
  vec3 e = vsub(p, vadd(a, vscale(d, t)));
  return vdot(e, e);
}
double oracle_point_triangle_distance2(const double *pp, const double *pa, const double *pb,
                                       const double *pc) {
  vec3 p = {pp[0], pp[1], pp[2]}, a = {pa[0], pa[1], pa[2]}, b = {pb[0], pb[1], pb[2]},
       c = {pc[0], pc[1], pc[2]};
  vec3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
  if (vdot(vcross(ab, ac), vcross(ab, ac)) < 1e-14 * 1e-14)
    return or_min(or_min(or_seg_d2(p, a, b), or_seg_d2(p, a, c)), or_seg_d2(p, b, c));
  double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
  if (d1 <= 0 && d2 <= 0)
    return vdot(ap, ap);
  vec3 bp = vsub(p, b);
  double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
  if (d3 >= 0 && d4 <= d3)
    return vdot(bp, bp);
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    vec3 e = vsub(p, vadd(a, vscale(ab, d1 / (d1 - d3))));
    return vdot(e, e);
  }
  vec3 cp = vsub(p, c);
  double d5 = vdot(ab, cp), d6 = vdot(ac, cp);
  if (d6 >= 0 && d5 <= d6)
    return vdot(cp, cp);
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    vec3 e = vsub(p, vadd(a, vscale(ac, d2 / (d2 - d6))));
    return vdot(e, e);
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    vec3 e = vsub(p, vadd(b, vscale(vsub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6)))));
    return vdot(e, e);
  }
  vec3 e = vsub(p, vadd(vadd(a, vscale(ab, vb / (va + vb + vc))), vscale(ac, vc / (va + vb + vc))));
  double r = vdot(e, e);
  return 0.0 < r ? r : 0.0;
}

/* src/cpu/verification.cpp:57-94, one row per (cluster, side, vertex), with
 * a brute-force scan of every face in place of the surface tree
 * (src/cpu/surface.h:85-95).  The minimum squared distance is the same value
 * (the tree skips only boxes strictly farther than the best so far); the
 * face the tree reports among several at exactly that distance depends on
 * its visiting order, so the normal test is evaluated for every such face. */
int oracle_verification_rows(int64_t nv, const double *x, const double *y, const double *z,
                             int64_t nf, const int32_t *f0, const int32_t *f1,
                             const int32_t *f2, const double *nx, const double *ny,
                             const double *nz, int64_t nc, const uint8_t *refl,
                             const double *rot, const double *trans, const double *pn,
                             const double *po, double verify_threshold, double box,
                             uint8_t *state, double *dist2) {
  double threshold = verify_threshold * box;
  for (int64_t c = 0; c < nc; ++c) {
    mat3 r, inv;
    for (int k = 0; k < 9; ++k)
      r.m[k] = rot[9 * c + k];
    inv = mtranspose(&r);
    vec3 t = {trans[3 * c], trans[3 * c + 1], trans[3 * c + 2]};
    vec3 n = {pn[3 * c], pn[3 * c + 1], pn[3 * c + 2]};
    for (int side = 0; side < 2; ++side) {
      const mat3 *m = side == 0 ? &r : &inv;
      for (int64_t i = 0; i < nv; ++i) {
        int64_t row = (2 * c + side) * nv + i;
        vec3 p = {x[i], y[i], z[i]};
        vec3 q = side == 0 ? vadd(mapply(&r, p), t) : mapply(&inv, vsub(p, t));
        state[row] = OR_VERIFY_EARLY;
        dist2[row] = NAN;
        if (vnorm(vsub(q, p)) < threshold * 2)
          continue;
        if (refl[c]) {
          double s = vdot(n, p) + po[c];
          if (side == 0 ? s >= -threshold : s <= threshold)
            continue;
        }
        double qq[3] = {q.x, q.y, q.z}, best = INFINITY;
        for (int64_t f = 0; f < nf; ++f) {
          double a[3] = {x[f0[f]], y[f0[f]], z[f0[f]]}, b[3] = {x[f1[f]], y[f1[f]], z[f1[f]]},
                 cc[3] = {x[f2[f]], y[f2[f]], z[f2[f]]};
          double d = oracle_point_triangle_distance2(qq, a, b, cc);
          if (d < best)
            best = d;
        }
        dist2[row] = best;
        if (best > threshold * threshold) {
          state[row] = OR_VERIFY_FAR;
          continue;
        }
        vec3 vn = mapply(m, (vec3){nx[i], ny[i], nz[i]});
        int pass = 0, failed = 0;
        for (int64_t f = 0; f < nf; ++f) {
          vec3 a = {x[f0[f]], y[f0[f]], z[f0[f]]}, b = {x[f1[f]], y[f1[f]], z[f1[f]]},
               cc = {x[f2[f]], y[f2[f]], z[f2[f]]};
          double pa[3] = {a.x, a.y, a.z}, pb[3] = {b.x, b.y, b.z}, pc[3] = {cc.x, cc.y, cc.z};
          if (oracle_point_triangle_distance2(qq, pa, pb, pc) != best)
            continue;
          vec3 fnrm = vunit(vcross(vsub(b, a), vsub(cc, a)));
          if (vdot(vn, fnrm) < 0.5)
            failed = 1;
          else
            pass = 1;
        }
        state[row] = pass && failed ? OR_VERIFY_TIE : (pass ? OR_VERIFY_PASS : OR_VERIFY_NORMAL);
      }
    }
  }
  return OR_OK;
}

/* src/cpu/verification.cpp:22-49: breadth-first components of the vertices
 * `ids` over the neighbour rings, seeded in the order of `ids`; each patch
 * sorted.  off: patches+1 (capacity n_ids+1), members: n_ids. */
int oracle_components(int64_t nv, const uint64_t *nbr_off, const int32_t *nbr_idx, int64_t n_ids,
                      const int32_t *ids, uint64_t *off, int32_t *members, int64_t *patches) {
  uint8_t *allowed = calloc((size_t)(nv > 0 ? nv : 1), 1);
  uint8_t *seen = calloc((size_t)(nv > 0 ? nv : 1), 1);
  if (!allowed || !seen) {
    free(allowed);
    free(seen);
    return fail(OR_ERUNTIME, "allocation failed");
  }
  for (int64_t k = 0; k < n_ids; ++k)
    allowed[ids[k]] = 1;
  int64_t np = 0, pos = 0;
  off[0] = 0;
  for (int64_t k = 0; k < n_ids; ++k) {
    int32_t seed = ids[k];
    if (seen[seed])
      continue;
    /* the queue is the patch itself: members[start .. tail) */
    int64_t start = pos, head = pos;
    members[pos++] = seed;
    seen[seed] = 1;
    while (head < pos) {
      int32_t v = members[head++];
      for (uint64_t j = nbr_off[v]; j < nbr_off[v + 1]; ++j) {
        int32_t w = nbr_idx[j];
        if (allowed[w] && !seen[w]) {
          seen[w] = 1;
          members[pos++] = w;
        }
      }
    }
    qsort(members + start, (size_t)(pos - start), sizeof(int32_t), cmp_i32);
    off[++np] = (uint64_t)pos;
  }
  *patches = np;
  free(allowed);
  free(seen);
  return OR_OK;
}

/* --------------------------------------------------- multi-threaded runs ---
 * The CPU baseline on all host cores: the single-thread functions above run on
 * independent units (pairing: ranges of sources; transformations: ranges of
 * pairs; batched mean-shift: ranges of seeds), chunks claimed in order by
 * `threads` POSIX threads, results concatenated in unit order, so every
 * output equals the single-thread call's (checked in tests/test_oracle.py).
 * Budgets are checked on the totals. */
#include <pthread.h>
#include <stdatomic.h>

typedef struct {
  int kind; /* 0 pairing, 1 transformations, 2 batched mean-shift */
  int64_t units, chunk;
  atomic_llong next;
  int rc;
  pthread_mutex_t lock;
  /* pairing */
  int64_t nv;
  const double *x, *y, *z, *kmin, *kmax, *feat;
  int64_t ns;
  const int32_t *sel;
  double box, t;
  const double *center;
  const int32_t *pa, *pb;
  int32_t **ca, **cb;   /* per chunk outputs */
  int64_t *cn;          /* per chunk counts */
  int32_t **cpair;
  uint8_t **ckind;
  double **clanes;
  int64_t *ccap;
  /* mean-shift */
  int dim, dim_rigid, dim_refl, iterations;
  const double *pts;
  const uint8_t *vkind, *seed_kind;
  const double *seeds;
  double h, cutoff;
  double *out, *weight;
} mt_job;

static void *mt_worker(void *arg) {
  mt_job *J = (mt_job *)arg;
  for (;;) {
    long long c = atomic_fetch_add(&J->next, 1);
    int64_t u0 = (int64_t)c * J->chunk;
    if (u0 >= J->units)
      break;
    int64_t u1 = u0 + J->chunk < J->units ? u0 + J->chunk : J->units;
    int rc = OR_OK;
    if (J->kind == 0) {
      /* at most one pair per (source, later sample) */
      int64_t cap = 0;
      for (int64_t a = u0; a < u1; ++a)
        cap += J->ns - 1 - a;
      J->ca[c] = (int32_t *)malloc((size_t)(cap ? cap : 1) * sizeof(int32_t));
      J->cb[c] = (int32_t *)malloc((size_t)(cap ? cap : 1) * sizeof(int32_t));
      if (!J->ca[c] || !J->cb[c])
        rc = fail(OR_EOVERFLOW, "out of host memory");
      else
        rc = pairing_range(J->nv, J->x, J->y, J->z, J->kmin, J->kmax, J->ns, J->sel, J->box,
                           J->t, UINT64_MAX, cap, J->ca[c], J->cb[c], &J->cn[c], u0, u1);
    } else if (J->kind == 1) {
      int64_t cap = 3 * (u1 - u0);
      J->ccap[c] = cap;
      J->cpair[c] = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
      J->ckind[c] = (uint8_t *)malloc((size_t)cap);
      J->clanes[c] = (double *)malloc((size_t)cap * 28 * sizeof(double));
      if (!J->cpair[c] || !J->ckind[c] || !J->clanes[c])
        rc = fail(OR_EOVERFLOW, "out of host memory");
      else
        rc = oracle_transformations(J->nv, J->x, J->y, J->z, J->feat, u1 - u0, J->pa + u0,
                                    J->pb + u0, J->box, J->center, UINT64_MAX, cap, J->cpair[c],
                                    J->ckind[c], J->clanes[c], &J->cn[c]);
    } else {
      int64_t m = u1 - u0;
      double *sd = (double *)malloc((size_t)(J->dim * m) * sizeof(double));
      double *so = (double *)malloc((size_t)(J->dim * m) * sizeof(double));
      if (!sd || !so) {
        rc = fail(OR_EOVERFLOW, "out of host memory");
      } else {
        for (int k = 0; k < J->dim; ++k)
          memcpy(sd + (int64_t)k * m, J->seeds + (int64_t)k * J->units + u0,
                 (size_t)m * sizeof(double));
        rc = oracle_meanshift_batched(J->nv, J->dim, J->pts, J->vkind, J->dim_rigid,
                                      J->dim_refl, m, sd, J->seed_kind + u0, J->h, J->cutoff,
                                      J->iterations, so, J->weight + u0);
        for (int k = 0; k < J->dim && rc == OR_OK; ++k)
          memcpy(J->out + (int64_t)k * J->units + u0, so + (int64_t)k * m,
                 (size_t)m * sizeof(double));
      }
      free(sd);
      free(so);
    }
    if (rc != OR_OK) {
      pthread_mutex_lock(&J->lock);
      J->rc = rc;
      pthread_mutex_unlock(&J->lock);
    }
  }
  return NULL;
}

static int mt_run(mt_job *J, int threads) {
  if (threads < 1)
    threads = 1;
  if (threads > 256)
    threads = 256;
  atomic_init(&J->next, 0);
  J->rc = OR_OK;
  pthread_mutex_init(&J->lock, NULL);
  pthread_t tid[256];
  int started = 0;
  for (int k = 1; k < threads; ++k)
    if (pthread_create(&tid[k], NULL, mt_worker, J) == 0)
      ++started;
    else
      break;
  mt_worker(J);
  for (int k = 1; k <= started; ++k)
    pthread_join(tid[k], NULL);
  pthread_mutex_destroy(&J->lock);
  return J->rc;
}

int oracle_pairing_mt(int threads, int64_t nv, const double *x, const double *y,
                      const double *z, const double *kmin, const double *kmax, int64_t ns,
                      const int32_t *sel, double box, double t, uint64_t max_pairs, int64_t cap,
                      int32_t *out_a, int32_t *out_b, int64_t *count) {
  if (ns < 0 || cap < 0)
    return fail(OR_EINVAL, "negative size");
  for (int64_t s = 0; s < ns; ++s)
    if (sel[s] < 0 || sel[s] >= nv)
      return fail(OR_EINVAL, "selected vertex index out of range");
  mt_job J;
  memset(&J, 0, sizeof J);
  J.kind = 0;
  J.units = ns;
  J.chunk = 64;
  J.nv = nv, J.x = x, J.y = y, J.z = z, J.kmin = kmin, J.kmax = kmax, J.ns = ns, J.sel = sel;
  J.box = box, J.t = t;
  int64_t nc = (ns + J.chunk - 1) / J.chunk;
  J.ca = (int32_t **)calloc((size_t)(nc ? nc : 1), sizeof(int32_t *));
  J.cb = (int32_t **)calloc((size_t)(nc ? nc : 1), sizeof(int32_t *));
  J.cn = (int64_t *)calloc((size_t)(nc ? nc : 1), sizeof(int64_t));
  int rc = J.ca && J.cb && J.cn ? mt_run(&J, threads) : fail(OR_EOVERFLOW, "out of host memory");
  int64_t c = 0;
  for (int64_t k = 0; k < nc && rc == OR_OK; ++k) {
    if ((uint64_t)(c + J.cn[k]) > max_pairs) {
      rc = fail(OR_EBUDGET, "Pair budget exceeded; lower --sample-ratio "
                            "or tighten --pair-threshold");
      break;
    }
    if (c + J.cn[k] > cap) {
      rc = fail(OR_EOVERFLOW, "pair output capacity exceeded");
      break;
    }
    memcpy(out_a + c, J.ca[k], (size_t)J.cn[k] * sizeof(int32_t));
    memcpy(out_b + c, J.cb[k], (size_t)J.cn[k] * sizeof(int32_t));
    c += J.cn[k];
  }
  for (int64_t k = 0; k < nc && J.ca; ++k) {
    free(J.ca[k]);
    free(J.cb[k]);
  }
  free(J.ca);
  free(J.cb);
  free(J.cn);
  if (rc == OR_OK)
    *count = c;
  return rc;
}

int oracle_transformations_mt(int threads, int64_t nv, const double *x, const double *y,
                              const double *z, const double *feat, int64_t np,
                              const int32_t *pa, const int32_t *pb, double box,
                              const double *center, uint64_t max_votes, int64_t cap,
                              int32_t *pair, uint8_t *kind, double *lanes, int64_t *count) {
  if (np < 0 || cap < 0)
    return fail(OR_EINVAL, "negative size");
  mt_job J;
  memset(&J, 0, sizeof J);
  J.kind = 1;
  J.units = np;
  J.chunk = 4096;
  J.nv = nv, J.x = x, J.y = y, J.z = z, J.feat = feat, J.pa = pa, J.pb = pb, J.box = box;
  J.center = center;
  int64_t nc = (np + J.chunk - 1) / J.chunk;
  J.cpair = (int32_t **)calloc((size_t)(nc ? nc : 1), sizeof(int32_t *));
  J.ckind = (uint8_t **)calloc((size_t)(nc ? nc : 1), sizeof(uint8_t *));
  J.clanes = (double **)calloc((size_t)(nc ? nc : 1), sizeof(double *));
  J.ccap = (int64_t *)calloc((size_t)(nc ? nc : 1), sizeof(int64_t));
  J.cn = (int64_t *)calloc((size_t)(nc ? nc : 1), sizeof(int64_t));
  int rc = J.cpair && J.ckind && J.clanes && J.ccap && J.cn
               ? mt_run(&J, threads)
               : fail(OR_EOVERFLOW, "out of host memory");
  int64_t c = 0;
  for (int64_t k = 0; k < nc && rc == OR_OK; ++k) {
    if ((uint64_t)(c + J.cn[k]) > max_votes) {
      rc = fail(OR_EBUDGET, "Vote budget exceeded; lower --sample-ratio or "
                            "tighten --pair-threshold");
      break;
    }
    if (c + J.cn[k] > cap) {
      rc = fail(OR_EOVERFLOW, "vote output capacity exceeded");
      break;
    }
    for (int64_t v = 0; v < J.cn[k]; ++v) {
      pair[c + v] = J.cpair[k][v] + (int32_t)(k * J.chunk);
      kind[c + v] = J.ckind[k][v];
    }
    for (int l = 0; l < 28; ++l)
      memcpy(lanes + (int64_t)l * cap + c, J.clanes[k] + (int64_t)l * J.ccap[k],
             (size_t)J.cn[k] * sizeof(double));
    c += J.cn[k];
  }
  for (int64_t k = 0; k < nc && J.cpair; ++k) {
    free(J.cpair[k]);
    free(J.ckind[k]);
    free(J.clanes[k]);
  }
  free(J.cpair);
  free(J.ckind);
  free(J.clanes);
  free(J.ccap);
  free(J.cn);
  if (rc == OR_OK)
    *count = c;
  return rc;
}

int oracle_meanshift_batched_mt(int threads, int64_t nv, int dim, const double *pts,
                                const uint8_t *kind, int dim_rigid, int dim_refl, int64_t ns,
                                const double *seeds, const uint8_t *seed_kind, double h,
                                double cutoff, int iterations, double *out, double *weight) {
  mt_job J;
  memset(&J, 0, sizeof J);
  J.kind = 2;
  J.units = ns;
  J.chunk = 16;
  J.nv = nv, J.dim = dim, J.pts = pts, J.vkind = kind, J.dim_rigid = dim_rigid;
  J.dim_refl = dim_refl, J.seeds = seeds, J.seed_kind = seed_kind, J.h = h, J.cutoff = cutoff;
  J.iterations = iterations, J.out = out, J.weight = weight;
  return mt_run(&J, threads);
}
