/*
 * CPU oracle for the transform-space voting path of the symmetry detector.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference CPU backend.  Only tests/, the smoke entry and
 * the CPU-baseline leg of the benchmark may load it, and only as the checker.
 * The product path (libmvs) never links, loads or calls anything in oracle/.
 *
 * Every function cites the reference source it restates (paths relative to the
 * reference repository).  Operation order and forms follow the reference so
 * that results are bit-identical with the reference CPU build on the same
 * machine (checked against oracle/_ref in tests/test_oracle.py).
 *
 * Pinning (what checks each function besides the compiled reference):
 *   or_mt64_*            C++ standard known answer (10000th draw of the
 *                        default-seeded mt19937_64 = 9981545732273789042).
 *   oracle_prepare       tetrahedron topology expectations
 *                        (tests/storage_backend.cpp:132-147), brute force.
 *   oracle_curvature     invariants (tests/cuda_curvature.cpp:22-25), planar
 *                        mesh gives all-invalid zeros, cylinder anisotropy.
 *   oracle_sample_vertices  closed-form exponential keys on small inputs.
 *   oracle_prune         strict-boundary fixture keeps 2/3
 *                        (tests/cuda_backend.cpp:338-355).
 *   oracle_pairing       known answers 495000 / 131328 / 190 / 1 / 0
 *                        (benchmarks/pairing.cpp, tests/cuda_pairing.cpp),
 *                        brute force.
 *   oracle_transformations  R R^T = I, det = +-1, R p + t = q, |n| = 1,
 *                        3 votes/pair mirror fixture, 2 votes/pair parallel
 *                        normals (tests/cuda_backend.cpp:33-143).
 *   oracle_clustering    4-mode synthetic votes give 4 clusters
 *                        (benchmarks/results/rtx4070_wsl_clustering.json).
 *   end to end           examples/{reflection,rigid,partial,single}.off counts 1898/4452, 1898/4328,
 *                        2378/5630, 458/1122 (benchmarks/results/
 *                        rtx4070_wsl_cli.json) and 4470/10356 at t=.15
 *                        (CUDA_STAGE3_RU.md:125).
 *   oracle_point_triangle_distance2  known answers (tests/geometry.cpp:8-20)
 *                        and the compiled reference on random and degenerate
 *                        triangles (the query set of
 *                        tests/cuda_verification.cpp:77-99).
 *   oracle_verification_rows, oracle_components  the compiled reference's
 *                        Backend::verification on the example meshes
 *                        (sides, patches, rms, coverage); mirror and half-turn
 *                        closed forms on exactly symmetric grids.
 *   oracle_meanshift_batched  has no reference counterpart (the reference's
 *                        seeds are sequential).  Pinned by what the
 *                        mathematics fixes (tests/test_oracle_meanshift.py):
 *                        converged modes are stationary points and local
 *                        maxima of the density whose profile derivative is
 *                        the weight (FP64 finite differences of the density
 *                        alone), the density never decreases along the
 *                        iterates, symmetric point sets keep a seed on the
 *                        centre or symmetry hyperplane exactly, points at
 *                        (1 -/+ 1e-9) c h are inside / outside the support,
 *                        and each kind sees only its own votes and lanes.
 */
#ifndef MVS_ORACLE_H
#define MVS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { OR_OK = 0, OR_EINVAL = 1, OR_EBUDGET = 2, OR_EOVERFLOW = 3, OR_ERUNTIME = 6 };
const char *oracle_last_error(void);

/* std::mt19937_64 with libstdc++'s state layout (x[312], index p). */
typedef struct {
  uint64_t x[312];
  uint64_t p;
} or_mt64;
void or_mt64_seed(or_mt64 *g, uint64_t seed);
uint64_t or_mt64_next(or_mt64 *g);

/* Edge record, same field order as symmetry::Edge (include/symmetry/pipeline.h:11-13). */
typedef struct {
  int32_t lo, hi, a, b, face0, face1;
} or_edge;

/* prepare() outputs; every array is caller-allocated with the stated capacity. */
typedef struct {
  int64_t face_count;         /* usable faces after degenerate rejection */
  int32_t *f0, *f1, *f2;      /* capacity nf */
  double *nx, *ny, *nz;       /* vertex normals, nv */
  double *fnx, *fny, *fnz;    /* face normals, capacity nf */
  double *weights;            /* area/3 per corner, nv */
  uint64_t *nbr_off;          /* nv+1 */
  int32_t *nbr_idx;           /* capacity 6*nf */
  int64_t nbr_nnz;
  uint64_t *inc_off;          /* nv+1 */
  int32_t *inc_idx;           /* capacity 3*nf */
  int64_t inc_nnz;
  or_edge *edges;             /* capacity 3*nf */
  int64_t edge_count;
  double box_size, center[3];
  int32_t rejected_faces, nonmanifold_edges, orientation_conflicts;
} or_prepared;

/* src/core/prepare.cpp:102-149 (prepare) and :13-58 (topology); subdivide=0. */
int oracle_prepare(int64_t nv, const double *x, const double *y, const double *z,
                   int64_t nf, const int32_t *f0, const int32_t *f1, const int32_t *f2,
                   or_prepared *out);

/* src/cpu/curvature.cpp:14-79.  feat is an 11*nv SoA (lane stride nv):
 * normal xyz, dmin xyz, dmax xyz, kmin, kmax. */
int oracle_curvature(int64_t nv, const double *x, const double *y, const double *z,
                     const double *nx, const double *ny, const double *nz,
                     const double *fnx, const double *fny, const double *fnz,
                     const uint64_t *nbr_off, const int32_t *nbr_idx,
                     int64_t ne, const or_edge *edges, int smoothing,
                     double *feat, uint8_t *valid);

/* src/core/sampling.cpp:5-24.  out capacity nv. */
int oracle_sample_vertices(int64_t nv, const double *weights, double ratio,
                           or_mt64 *rng, int32_t *out, int64_t *count);

/* src/cpu/selection.cpp:21-27 (prune predicate and order-preserving compaction). */
int oracle_prune(int64_t nv, const double *kmin, const double *kmax,
                 const uint8_t *valid, int64_t ns, const int32_t *sample,
                 double prune, int32_t *out, int64_t *count);

/* src/cpu/pairing.cpp:14-35.  Output capacity = cap (>= min(max_pairs, all)). */
int oracle_pairing(int64_t nv, const double *x, const double *y, const double *z,
                   const double *kmin, const double *kmax, int64_t ns,
                   const int32_t *selected, double box, double threshold,
                   uint64_t max_pairs, int64_t cap, int32_t *out_a, int32_t *out_b,
                   int64_t *count);

/* src/cpu/transformations.cpp:17-73.  lanes is a 28*cap SoA (stride cap):
 * rotation[9], translation[3], plane_normal[3], plane_offset, descriptor[12]. */
int oracle_transformations(int64_t nv, const double *x, const double *y,
                           const double *z, const double *feat, int64_t np,
                           const int32_t *pa, const int32_t *pb, double box,
                           const double *center, uint64_t max_votes, int64_t cap,
                           int32_t *pair, uint8_t *kind, double *lanes,
                           int64_t *count);

/* src/cpu/clustering.cpp:32-208 (sequential-seed Epanechnikov mean-shift,
 * merge, winners, extraction, stable sort, truncation).
 * desc: 12*nv SoA (stride nv); rotation: 9*nv SoA (stride nv).
 * Cluster outputs have capacity max_clusters; members capacity nv. */
int oracle_clustering(int64_t nv, const uint8_t *kind, const double *desc,
                      const double *rotation, double box, const double *center,
                      double bandwidth, int max_iterations, int max_clusters,
                      or_mt64 *rng, int32_t *n_clusters, uint8_t *c_reflection,
                      double *c_rotation, double *c_translation,
                      double *c_plane_normal, double *c_plane_offset,
                      uint64_t *member_off, int32_t *members,
                      int32_t *limited_iterations);

/* Batched-seed mean-shift (no reference counterpart).  Each seed s runs
 * exactly `iterations` Gaussian-kernel updates over all votes of its kind,
 * with the Gaussian truncated at cutoff*h and shifted to reach zero there
 * (continuous kernel):
 *   w_j = exp(-d2_j / (2 h^2)) - exp(-cutoff^2 / 2) for d2_j < (cutoff*h)^2
 *   mean <- mean + sum_j w_j (x_j - mean) / sum_j w_j  (unchanged if no support)
 * points: dim*nv SoA (stride nv), kinds per vote; seeds: dim*ns SoA holding the
 * start positions and seed_kind per seed.  Rigid votes use the first
 * dim_rigid lanes, reflection votes the first dim_reflection lanes.
 * out: dim*ns SoA, plus the final kernel weight sum per seed. */
int oracle_meanshift_batched(int64_t nv, int dim, const double *points,
                             const uint8_t *kind, int dim_rigid,
                             int dim_reflection, int64_t ns, const double *seeds,
                             const uint8_t *seed_kind, double h, double cutoff,
                             int iterations, double *out, double *weight);

/* The same three computations on `threads` host threads (units: sources,
 * pairs, seeds), outputs identical to the single-thread calls'. */
int oracle_pairing_mt(int threads, int64_t nv, const double *x, const double *y,
                      const double *z, const double *kmin, const double *kmax, int64_t ns,
                      const int32_t *sel, double box, double t, uint64_t max_pairs, int64_t cap,
                      int32_t *out_a, int32_t *out_b, int64_t *count);
int oracle_transformations_mt(int threads, int64_t nv, const double *x, const double *y,
                              const double *z, const double *feat, int64_t np,
                              const int32_t *pa, const int32_t *pb, double box,
                              const double *center, uint64_t max_votes, int64_t cap,
                              int32_t *pair, uint8_t *kind, double *lanes, int64_t *count);
int oracle_meanshift_batched_mt(int threads, int64_t nv, int dim, const double *points,
                                const uint8_t *kind, int dim_rigid, int dim_reflection,
                                int64_t ns, const double *seeds, const uint8_t *seed_kind,
                                double h, double cutoff, int iterations, double *out,
                                double *weight);

/* src/core/math.cpp:72-112. */
double oracle_point_triangle_distance2(const double *p, const double *a, const double *b,
                                       const double *c);

/* Row states of oracle_verification_rows. */
enum {
  OR_VERIFY_EARLY = 0,   /* moved less than 2 tau, or on the plane's near side */
  OR_VERIFY_FAR = 1,     /* nearest face farther than tau */
  OR_VERIFY_NORMAL = 2,  /* normal test fails for every face at the minimum */
  OR_VERIFY_PASS = 3,    /* accepted: passes for every face at the minimum */
  OR_VERIFY_TIE = 4      /* faces at the minimum disagree: the reference's
                            visiting order decides */
};
/* src/cpu/verification.cpp:57-94 per row (cluster c, side s, vertex i) at
 * (2c+s)*nv + i, brute-force nearest face: state and the minimum squared
 * distance (NaN for OR_VERIFY_EARLY).  rot: 9*nc row-major, trans/pn: 3*nc. */
int oracle_verification_rows(int64_t nv, const double *x, const double *y, const double *z,
                             int64_t nf, const int32_t *f0, const int32_t *f1,
                             const int32_t *f2, const double *nx, const double *ny,
                             const double *nz, int64_t nc, const uint8_t *refl,
                             const double *rot, const double *trans, const double *pn,
                             const double *po, double verify_threshold, double box,
                             uint8_t *state, double *dist2);
/* src/cpu/verification.cpp:22-49: components of `ids` (breadth-first, seeds
 * in the order of ids, each patch sorted). */
int oracle_components(int64_t nv, const uint64_t *nbr_off, const int32_t *nbr_idx, int64_t n_ids,
                      const int32_t *ids, uint64_t *off, int32_t *members, int64_t *patches);

#ifdef __cplusplus
}
#endif
#endif
