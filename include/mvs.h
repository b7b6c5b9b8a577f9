/*
 * libmvs — MI355X (gfx950) transform-space voting core of the partial-symmetry
 * detector.  C ABI: plain pointers and sizes, no C++ or torch types.
 *
 * Two layers:
 *   1. Host layer (mvs_curvature, mvs_prune, mvs_pairing, mvs_transformations,
 *      mvs_clustering, mvs_verification): borrowed HOST views in, host
 *      results out, synchronous.
 *      Each call is the drop-in replacement of one method of the reference's
 *      `symmetry::Backend` (include/symmetry/pipeline.h:54-69) with the same
 *      argument meaning, output order and error classes.
 *   2. Device layer (mvs_dev_*): DEVICE pointers in HBM, ordered on the
 *      session's stream; votes never leave HBM.  This is the fast path.
 *
 * Status codes map to the reference's exception classes:
 *   MVS_EINVAL    std::invalid_argument (contract / size violations)
 *   MVS_EBUDGET   std::runtime_error whose message contains "budget"
 *   MVS_EOVERFLOW std::overflow_error (int32 / size overflow)
 *   MVS_EHIP      std::runtime_error "HIP: ..." (device errors)
 *   MVS_ENODEV    std::runtime_error "HIP: no available device"
 *   MVS_ERUNTIME  other std::runtime_error (mean-shift budgets, ...)
 * mvs_last_error() returns the message of the calling thread's last failure.
 *
 * Sessions are not thread-safe; use one per host thread.  Calls on one
 * session are serialised on its stream.
 */
#ifndef MVS_H
#define MVS_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  MVS_OK = 0,
  MVS_EINVAL = 1,
  MVS_EBUDGET = 2,
  MVS_EOVERFLOW = 3,
  MVS_EHIP = 4,
  MVS_ENODEV = 5,
  MVS_ERUNTIME = 6
};

/* Arithmetic of the pairing predicate, frames, votes and mean-shift. */
enum {
  MVS_FP64 = 0,          /* plain FP64 throughout (reference arithmetic) */
  MVS_FP32_GUARDED = 1   /* FP32 with FP64 recheck inside guard bands */
};

typedef struct mvs_session mvs_session;

/* Edge record with the field order of symmetry::Edge
 * (include/symmetry/pipeline.h:11-13). */
typedef struct {
  int32_t lo, hi, a, b, face0, face1;
} mvs_edge;

/* std::mt19937_64 state in libstdc++ layout (x[312], index p); the state the
 * reference's clustering consumes (src/cpu/clustering.cpp:51). */
typedef struct {
  uint64_t x[312];
  uint64_t p;
} mvs_mt64;

/* SoA point view: x, y, z arrays of `count` doubles (PointsView,
 * include/symmetry/storage.h:49-57). */
typedef struct {
  uint64_t count;
  const double *x, *y, *z;
} mvs_points;

/* Topology view (TopologyView, include/symmetry/pipeline.h:14-20).
 * Only normals, face_normals, neighbors and edges are read by curvature. */
typedef struct {
  mvs_points normals;
  mvs_points face_normals;
  const uint64_t *neighbor_offsets; /* count+1, first 0 */
  const int32_t *neighbor_indices;
  uint64_t edge_count;
  const mvs_edge *edges;
} mvs_topology;

/* Feature lanes in FeaturesView order (include/symmetry/storage.h:213-225):
 * 0-2 normal xyz, 3-5 dmin xyz, 6-8 dmax xyz, 9 kmin, 10 kmax; plus valid. */
typedef struct {
  uint64_t count;
  const double *lane[11];
  const uint8_t *valid;
} mvs_features;
typedef struct {
  double *lane[11];
  uint8_t *valid;
} mvs_features_out;

/* Mesh preparation outputs (prepare(), src/core/prepare.cpp:102-149, with the
 * topology of :13-58).  Caller-allocated arrays with the stated capacities;
 * counts are written back. */
typedef struct {
  int64_t face_count;       /* usable faces after degenerate rejection */
  int32_t *f0, *f1, *f2;    /* capacity nf */
  double *nx, *ny, *nz;     /* vertex normals, nv */
  double *fnx, *fny, *fnz;  /* face normals, capacity nf */
  double *weights;          /* vertex area weights, nv */
  uint64_t *nbr_off;        /* nv+1 */
  int32_t *nbr_idx;         /* capacity 6*nf */
  int64_t nbr_nnz;
  uint64_t *inc_off;        /* nv+1 */
  int32_t *inc_idx;         /* capacity 3*nf */
  int64_t inc_nnz;
  mvs_edge *edges;          /* capacity 3*nf */
  int64_t edge_count;
  double box_size, center[3];
  int32_t rejected_faces, nonmanifold_edges, orientation_conflicts;
} mvs_prepared;

/* prepare() (src/core/prepare.cpp:102-149) on the host: validation,
 * degenerate-face rejection, face/vertex normals, area weights, sorted rings,
 * incident faces and the (lo, hi)-ordered edge list, box size and centre.
 * Errors carry the reference's messages (MVS_ERUNTIME). */
int mvs_prepare(int64_t nv, const double *x, const double *y, const double *z, int64_t nf,
                const int32_t *f0, const int32_t *f1, const int32_t *f2, mvs_prepared *out);

/* prepare() on the device: the same outputs as mvs_prepare, bit for bit, but
 * every array pointer in *out is a DEVICE pointer (same capacities) and the
 * mesh is also written to xyz, 3*nv device doubles (x, then y, then z), so
 * curvature and the stages after it start from HBM.  The mesh itself comes
 * from host arrays, as the reference's Mesh does; counts, box size, centre
 * and diagnostics come back in *out on the host.  Same errors and messages
 * as mvs_prepare (src/core/prepare.cpp:102-149); at most 2^32/3 faces.
 * Work is queued on the session stream; the call reads back two counts. */
int mvs_dev_prepare(mvs_session *s, int64_t nv, const double *x, const double *y,
                    const double *z, int64_t nf, const int32_t *f0, const int32_t *f1,
                    const int32_t *f2, double *xyz, mvs_prepared *out);

/* std::mt19937_64 on a caller-held state (libstdc++ layout). */
void mvs_mt64_seed(mvs_mt64 *g, uint64_t seed);
uint64_t mvs_mt64_next(mvs_mt64 *g);

/* Sampling step of Backend::select (src/core/sampling.cpp:5-24): weighted
 * sampling without replacement by exponential keys, one draw per
 * positive-weight vertex; ceil(nv*ratio) ids (at most the positive-weight
 * count), ascending.  out capacity nv.  Runs on the host like the reference. */
int mvs_sample_vertices(uint64_t nv, const double *weights, double ratio, mvs_mt64 *rng,
                        int32_t *out, uint64_t *count);

/* ------------------------------------------------------------- sessions */
int mvs_session_create(int device, mvs_session **out);
void mvs_session_destroy(mvs_session *s);
const char *mvs_last_error(void);
/* Stream the device layer runs on (default: a non-blocking stream the
 * session owns).  Pass a hipStream_t: 0 is the null stream (ordered against
 * every blocking stream, torch's default stream among them); MVS_OWN_STREAM
 * selects the session's own stream again.  Work on the session stream is
 * ordered only against that stream: a caller whose buffers are produced or
 * recycled on another stream passes that stream here. */
#define MVS_OWN_STREAM ((void *)(intptr_t)-1)
int mvs_session_set_stream(mvs_session *s, void *hip_stream);
int mvs_session_sync(mvs_session *s);
/* Stream-ordered device-to-device copy (to read session-owned outputs). */
int mvs_dev_copy(mvs_session *s, void *dst, const void *src, uint64_t bytes);
/* Per-launch kernel timing with HIP events on the session stream (clears
 * the accumulated times).  mvs_dev_timing returns the accumulated device
 * time (ms) and launch count of one kernel name, or of every kernel when
 * kernel is NULL. */
int mvs_session_enable_timing(mvs_session *s, int enable);
/* Restricts timing to the comma-separated kernel names (NULL or "" = all). */
int mvs_session_timing_filter(mvs_session *s, const char *names);
int mvs_dev_timing(mvs_session *s, const char *kernel, double *ms, uint64_t *launches);
/* Host <-> device copies of a timed session are timed too, under the names
 * "copy_h2d" and "copy_d2h"; a NULL kernel sums kernels only.
 *
 * Device memory and copy counters of the calling thread's libmvs calls (all
 * sessions): bytes its device buffers hold now, their peak since the last
 * reset, and bytes copied host to device and device to host since then.  The
 * counterpart of the CUDA backend's per-stage profile (src/cuda/profile.cpp:
 * 61-93, peak_buffer_bytes, h2d_bytes, d2h_bytes; all_buffers_released is
 * live_bytes back at its value before the stage).  Reset sets the peak to
 * the live bytes and the copy counts to 0. */
typedef struct {
  uint64_t live_bytes, peak_bytes, h2d_bytes, d2h_bytes;
} mvs_memory_metrics;
int mvs_thread_metrics(mvs_memory_metrics *out);
void mvs_thread_metrics_reset(void);

/* ------------------------------------------------------------ host layer */

/* Backend::curvature (pipeline.h:57; src/cpu/curvature.cpp:14-79).
 * Reads mesh vertices and the topology; writes 11 lanes + valid. */
int mvs_curvature(mvs_session *s, const mvs_points *vertices, const mvs_topology *topo,
                  int smoothing, const mvs_features_out *out);

/* Prune step of Backend::select (pipeline.h:58-59; src/cpu/selection.cpp:21-27):
 * keeps sample ids with valid && max(|kmin|,|kmax|)>1e-10 &&
 * |kmin-kmax|/max > 1-prune, order preserved.  out capacity: sample_count. */
int mvs_prune(mvs_session *s, const mvs_features *f, uint64_t sample_count,
              const int32_t *sample, double prune, int32_t *out, uint64_t *count);
/* mvs_prune on device arrays: feat is the 11-lane SoA of mvs_dev_curvature
 * (stride n; kmin lane 9, kmax lane 10), valid n bytes, sample ns device ids;
 * the kept ids go to out (device, capacity ns) in order, their count to
 * *count (host). */
int mvs_dev_prune(mvs_session *s, uint64_t n, const double *feat, const uint8_t *valid,
                  uint64_t sample_count, const int32_t *sample, double prune, int32_t *out,
                  uint64_t *count);

/* Backend::pairing (pipeline.h:60-61; src/cpu/pairing.cpp:14-35).  Computes
 * the pair set in CPU order and keeps it on the device; *count receives P.
 * Fetch with mvs_pairing_fetch.  threshold in (0,1], box > 0 finite. */
int mvs_pairing(mvs_session *s, const mvs_points *vertices, const mvs_features *f,
                uint64_t selected_count, const int32_t *selected, double box,
                double threshold, uint64_t max_pairs, int precision, uint64_t *count);
int mvs_pairing_fetch(mvs_session *s, int32_t *a, int32_t *b);

/* Backend::transformations (pipeline.h:62-63; src/cpu/transformations.cpp:17-73).
 * *count receives V; mvs_transformations_fetch writes pair ids, kinds
 * (0 rigid, 1 reflection) and 28 lanes: rotation[9], translation[3],
 * plane_normal[3], plane_offset, descriptor[12] (storage.h:329-348). */
int mvs_transformations(mvs_session *s, const mvs_points *vertices, const mvs_features *f,
                        uint64_t pair_count, const int32_t *a, const int32_t *b,
                        double box, const double center[3], uint64_t max_votes,
                        uint64_t *count);
int mvs_transformations_fetch(mvs_session *s, int32_t *pair, uint8_t *kind,
                              double *const lanes[28]);

/* Backend::clustering (pipeline.h:64-65; src/cpu/clustering.cpp:32-208):
 * sequential-seed Epanechnikov mean-shift with the reference's seeding,
 * merge, winner and extraction rules.  desc: 12 lanes, rotation: 9 lanes
 * (used for the degenerate-frame fallback).  rng is consumed exactly as the
 * reference consumes it.  Results: *n_clusters records and *n_members
 * member ids; fetch with mvs_clustering_fetch. */
int mvs_clustering(mvs_session *s, uint64_t vote_count, const uint8_t *kind,
                   const double *const desc[12], const double *const rotation[9],
                   double box, const double center[3], double bandwidth,
                   int max_iterations, int max_clusters, mvs_mt64 *rng,
                   uint64_t *n_clusters, uint64_t *n_members, int32_t *limited_iterations);
int mvs_clustering_fetch(mvs_session *s, uint8_t *reflection, double *rotation9,
                         double *translation3, double *plane_normal3, double *plane_offset,
                         uint64_t *member_offsets, int32_t *members);

/* Totals of the last mvs_verification call, summed over clusters. */
typedef struct {
  uint64_t side_a, side_b;        /* accepted vertices */
  uint64_t patches_a, patches_b;  /* patches */
} mvs_verification_counts;

/* One side of a Verification (include/symmetry/storage.h:371-374): the side
 * as a CSR over clusters (side_offsets: clusters+1, side: side total), and
 * its patches as a grouped CSR (patch_groups: clusters+1 patch offsets,
 * patch_offsets: patches+1 member offsets, patch_members: side total).
 * NULL members are skipped by the fetch. */
typedef struct {
  uint64_t *side_offsets;
  int32_t *side;
  uint64_t *patch_groups;
  uint64_t *patch_offsets;
  int32_t *patch_members;
} mvs_verification_side;

/* Backend::verification (pipeline.h:67-68; src/cpu/verification.cpp:50-110).
 * For every cluster: side A holds the vertices p whose forward image R p + t,
 * side B those whose inverse image R^T (p - t), moves at least 2 tau, lies on
 * the far side of a reflection plane by more than tau, is within tau of the
 * mesh surface and carries the vertex normal to within 60 degrees of the
 * nearest face's normal; tau = verify_threshold * box.  Patches are the
 * connected components of each side over the one-ring graph, ordered by
 * their smallest vertex, members ascending (verification.cpp:22-49); the
 * neighbour rings must be symmetric, as prepare() produces them.  rms and
 * coverage follow verification.cpp:95-96.  The surface search reproduces the
 * reference's tree and visiting order, so the nearest face, and with it every
 * set, equals the reference's; the squared distances are bit-identical.
 * vertices/normals: nv points; faces: face_count triangles (f0, f1, f2);
 * weights: nv vertex areas; neighbor_offsets: nv+1.  Clusters: cluster_count
 * records (reflection flag, rotation row-major 9, translation 3, plane
 * normal 3, plane offset).  Fetch with mvs_verification_fetch. */
int mvs_verification(mvs_session *s, const mvs_points *vertices, uint64_t face_count,
                     const int32_t *f0, const int32_t *f1, const int32_t *f2,
                     const mvs_points *normals, const double *weights,
                     const uint64_t *neighbor_offsets, const int32_t *neighbor_indices,
                     uint64_t cluster_count, const uint8_t *reflection, const double *rotation9,
                     const double *translation3, const double *plane_normal3,
                     const double *plane_offset, double verify_threshold, double box,
                     mvs_verification_counts *counts);
/* The same verification in two steps, for callers that keep the mesh in HBM
 * and verify several cluster sets against one surface:
 * mvs_verification_surface builds the surface tree of the mesh (as
 * mvs_verification does) and keeps it on the session;
 * mvs_dev_verification then runs against it with DEVICE inputs: pts6 holds
 * 6*nv doubles (x, y, z, then the vertex normal x, y, z, each nv long) of the
 * same vertices, neighbor_offsets (nv+1) and neighbor_indices the one-ring
 * CSR; weights and the cluster records are HOST arrays.  The caller
 * guarantees the device CSR is well formed (offsets non-decreasing from 0,
 * indices below nv).  Results and fetch as for mvs_verification. */
int mvs_verification_surface(mvs_session *s, const mvs_points *vertices, uint64_t face_count,
                             const int32_t *f0, const int32_t *f1, const int32_t *f2);
int mvs_dev_verification(mvs_session *s, uint64_t nv, const double *pts6,
                         const uint64_t *neighbor_offsets, const int32_t *neighbor_indices,
                         const double *weights, uint64_t cluster_count, const uint8_t *reflection,
                         const double *rotation9, const double *translation3,
                         const double *plane_normal3, const double *plane_offset,
                         double verify_threshold, double box, mvs_verification_counts *counts);
/* Copies the last verification out: both sides, coverage and rms
 * (cluster_count each).  Any pointer may be NULL. */
int mvs_verification_fetch(mvs_session *s, const mvs_verification_side *side_a,
                           const mvs_verification_side *side_b, double *coverage, double *rms);
/* Per-row results of the last verification, 2*cluster_count*nv rows ordered
 * (cluster, side, vertex): accepted 0/1 and the nearest squared distance
 * (NaN for rows rejected before the surface search).  Either may be NULL. */
int mvs_verification_rows(mvs_session *s, uint8_t *accepted, double *dist2);
/* Work of the last verification, summed over all rows: out[0] triangle
 * distance tests, out[1] tree nodes visited, out[2] FP64 additions,
 * subtractions, multiplications and divisions of the box and triangle
 * distances (the path each test took); then host-clock microseconds of
 * out[3] the device phase (launches to the segment sizes), out[4] the
 * download of the accepted rows, out[5] the assembly of sides and patches. */
int mvs_verification_stats(mvs_session *s, uint64_t out[6]);

/* ---------------------------------------------------------- device layer
 * All pointers below are device pointers.  Work is ordered on the session
 * stream; the call returns once the launches are queued unless it must read
 * a count back (documented per call). */

/* Curvature on device-resident inputs.  feat: 11*n SoA (stride n), valid: n.
 * precision MVS_FP64 only for this stage (frames are reused by both paths). */
int mvs_dev_curvature(mvs_session *s, uint64_t n, const double *x, const double *y,
                      const double *z, const double *normals3 /* 3*n SoA */,
                      uint64_t faces, const double *face_normals3 /* 3*faces SoA */,
                      const uint64_t *nbr_off, const int32_t *nbr_idx, uint64_t edges,
                      const mvs_edge *edge, int smoothing, double *feat, uint8_t *valid);

/* Pair planning on device-resident features (feat: 11*n SoA).  Gathers the
 * selected samples, sorts the curvature signatures, builds the
 * (kmin bucket, kmax) index, counts pairs per source and returns the total
 * (synchronises once to read it).  Budget errors are raised before any
 * output is written.  center is the model centre: the FP32 path works in
 * centre-relative coordinates.  The selected list must stay valid until
 * mvs_dev_pair_fill returns. */
int mvs_dev_pair_count(mvs_session *s, uint64_t n, const double *x, const double *y,
                       const double *z, const double *feat, uint64_t selected_count,
                       const int32_t *selected, double box, const double center[3],
                       double threshold, uint64_t max_pairs, int precision, uint64_t *pairs);
/* Writes the planned pairs in CPU order: a/b hold SELECTED ORDINALS when
 * ordinals != 0, else mesh vertex ids.  Capacity: the count returned above. */
int mvs_dev_pair_fill(mvs_session *s, int32_t *a, int32_t *b, int ordinals);
/* Statistics of the last pair count: kmin-window candidates, candidates
 * after the (kmin bucket, kmax) index, FP64 rechecks of the FP32 screen. */
int mvs_dev_pair_stats(mvs_session *s, uint64_t *window_candidates, uint64_t *index_candidates,
                       uint64_t *fp64_rechecks);

/* Compact vote emission for pairs given as selected ordinals (a < b).
 * Counts reflection votes (synchronises once) unless max_votes == 0 is
 * passed to mvs_dev_vote_emit directly after a count. Layout (SoA, s = 8
 * bytes for MVS_FP64, 4 bytes for MVS_FP32_GUARDED):
 *   rigid stream, 2 per pair (slot 2k + (sign<0)): q[4] (w,x,y,z, w >= 0),
 *       t[3], pair id (int32, = slot / 2; written only when rigid_pair is
 *       not NULL)
 *   reflection stream, in pair order: n[3], d, pair id (int32)
 * Global vote order (reference order) is recoverable: pair k's votes start
 * at 2k + (reflections before k); its reflection vote, if any, comes first. */
int mvs_dev_vote_count(mvs_session *s, uint64_t pair_count, const int32_t *a,
                       const int32_t *b, int precision, uint64_t *reflections);
int mvs_dev_vote_emit(mvs_session *s, uint64_t pair_count, const int32_t *a,
                      const int32_t *b, double box, const double center[3], int precision,
                      void *rigid_q4, void *rigid_t3, int32_t *rigid_pair, void *refl_n3,
                      void *refl_d, int32_t *refl_pair);

/* Fused pairing + vote emission on device-resident features: count, budget
 * check (one synchronisation for the pair total), placement, mirror
 * offsets and compact emission in one call, outputs in session-owned device
 * buffers valid until the next mvs_dev_pairvote on this session.  The
 * mirror stream has capacity pair_count (its SoA stride); the mirror total
 * is read back at the end.  Budgets: pairs > max_pairs or votes > max_votes
 * raise MVS_EBUDGET. */
typedef struct {
  uint64_t pairs, reflections;
  const int32_t *a, *b;             /* pair ordinals (selected positions) */
  const void *rigid_q4, *rigid_t3;  /* 4 / 3 lanes, stride 2*pairs */
  const int32_t *rigid_pair;        /* NULL: rigid vote v belongs to pair v / 2 */
  const void *refl_n3, *refl_d;     /* 3 / 1 lanes, stride refl_stride */
  const int32_t *refl_pair;
  uint64_t refl_stride;
  int precision;
} mvs_pairvote_result;
int mvs_dev_pairvote(mvs_session *s, uint64_t n, const double *x, const double *y,
                     const double *z, const double *feat, uint64_t selected_count,
                     const int32_t *selected, double box, const double center[3],
                     double threshold, uint64_t max_pairs, uint64_t max_votes, int precision,
                     mvs_pairvote_result *out);

/* Batched mean-shift over the votes of the last mvs_dev_pairvote:
 * `seeds` seed votes split between rigid and mirror votes in proportion to
 * their counts.  Within each stream of n votes, sorted by grid cell, seed j
 * of k is drawn from the j-th of k equal strata of the sorted order by a
 * splitmix64 hash of (seed_value, j): seeds are distinct and spread over the
 * occupied cells; seed_votes lists their stream indices in that (cell)
 * order.  Each seed is iterated up to `iterations` times in its own point
 * set (see mvs_dev_ms_batched).  Grid, seeds and queries are built on the
 * device; the call synchronises once at the end (non-finite points raise
 * MVS_EINVAL).  Outputs live in session-owned buffers; want_stats != 0 also
 * returns the visit counts. */
typedef struct {
  uint64_t rigid_seeds, refl_seeds;
  const int32_t *seed_votes;          /* rigid seeds then mirror seeds (stream indices) */
  const void *rigid_modes, *rigid_weight;  /* 6 x rigid_seeds SoA, rigid_seeds */
  const void *refl_modes, *refl_weight;    /* 4 x refl_seeds SoA, refl_seeds */
  uint64_t candidates, in_support;
  int precision;
  /* want_stats only: updates computed over all seeds (a seed stops once an
   * update leaves it unchanged), and the visit counts of each kind */
  uint64_t updates;
  uint64_t rigid_candidates, rigid_in_support, refl_candidates, refl_in_support;
} mvs_meanshift_result;
int mvs_dev_meanshift_votes(mvs_session *s, uint64_t seeds, uint64_t seed_value, int iterations,
                            double h, double cutoff, int want_stats, mvs_meanshift_result *out);

/* Batches of meshes (no reference counterpart: the reference runs one mesh
 * per detect call, src/core/pipeline.cpp:5-52).  The meshes' vertices and
 * features are concatenated (x, y, z, feat exactly as for one mesh, n = all
 * vertices); `selected` (device) lists every mesh's sample vertex ids as ids
 * into the concatenation, mesh m owning samples [sample_offsets[m],
 * sample_offsets[m + 1]).  box (meshes) and center (3 * meshes) describe each
 * mesh; the three metadata arrays are HOST arrays.  Pairs form inside each
 * mesh only.  For every mesh the pairs, their order and their compact votes
 * equal those of mvs_dev_pairvote on that mesh alone, with batch-wide
 * ordinals (mesh m's local ordinal = ordinal - sample_offsets[m]).  Budgets
 * apply to the whole batch.  Outputs: the fused result over all meshes, in
 * mesh order, and per-mesh starts of the pair list and of the mirror stream
 * (device arrays of meshes + 1). */
typedef struct {
  mvs_pairvote_result all;
  uint32_t meshes;
  const uint64_t *pair_offsets;  /* device, meshes + 1 */
  const uint64_t *refl_offsets;  /* device, meshes + 1 */
} mvs_batch_pairvote_result;
int mvs_dev_batch_pairvote(mvs_session *s, uint64_t n, const double *x, const double *y,
                           const double *z, const double *feat, uint32_t meshes,
                           const uint64_t *sample_offsets, const int32_t *selected,
                           const double *box, const double *center, double threshold,
                           uint64_t max_pairs, uint64_t max_votes, int precision,
                           mvs_batch_pairvote_result *out);

/* Batched mean-shift over the votes of the last mvs_dev_batch_pairvote:
 * mesh m draws ns = min(V_m, ceil(seeds_per_sample * S_m)) seeds (S_m its
 * samples, V_m its votes), split between its rigid and mirror votes as
 * mvs_dev_meanshift_votes splits `seeds`; each mesh's seeds, grids and
 * iterations equal that one-mesh call with seeds = ns under the same launch
 * regime (the seed count of the call, here all meshes' seeds, against the
 * session's wide-seed threshold decides whether a seed gets a block and a
 * finer grid, so compare with a one-mesh session whose threshold puts its
 * seeds in the same regime).  Outputs: modes and
 * weights in the layout of mvs_meanshift_result with all.rigid_seeds /
 * all.refl_seeds the capacities (SoA strides), mesh m's seeds at
 * [offsets[m], offsets[m + 1]) of each stream (device arrays of meshes + 1);
 * seed vote ids: rigid at [0, capacity), mirror after them, as indices into
 * the batch's rigid / mirror streams. */
typedef struct {
  mvs_meanshift_result all;
  uint32_t meshes;
  const uint64_t *rigid_seed_offsets;  /* device, meshes + 1 */
  const uint64_t *refl_seed_offsets;   /* device, meshes + 1 */
} mvs_batch_meanshift_result;
int mvs_dev_batch_meanshift(mvs_session *s, double seeds_per_sample, uint64_t seed_value,
                            int iterations, double h, double cutoff, int want_stats,
                            mvs_batch_meanshift_result *out);

/* One batch step (mvs_dev_batch_pairvote then mvs_dev_batch_meanshift with
 * these arguments, want_stats off) as a HIP graph owned by the session.
 * mvs_batch_graph_capture runs the step once (results in *pv, *ms, whose
 * buffers every replay rewrites) while recording each count the step reads
 * back, then captures the step on the session's own stream with those
 * counts frozen; each read-back site becomes a device check.
 * mvs_batch_graph_launch replays it with no host synchronisation, ordered
 * after the session stream's queued work and before its later work.
 * mvs_batch_graph_check synchronises and raises MVS_ERUNTIME when a replay
 * read counts other than the captured ones.  Replays are for inputs with the
 * captured counts (the same meshes and curvatures, as when a step is timed or
 * repeated): capture again after inputs change.  Writes are clamped to the
 * captured capacities, but a replay over inputs with other counts is not a
 * supported use.  Any call that grows a session buffer, and any new batch
 * call, drops the graph. */
int mvs_batch_graph_capture(mvs_session *s, uint64_t n, const double *x, const double *y,
                            const double *z, const double *feat, uint32_t meshes,
                            const uint64_t *sample_offsets, const int32_t *selected,
                            const double *box, const double *center, double threshold,
                            uint64_t max_pairs, uint64_t max_votes, int precision,
                            double seeds_per_sample, uint64_t seed_value, int iterations,
                            double h, double cutoff, mvs_batch_pairvote_result *pv,
                            mvs_batch_meanshift_result *ms);
int mvs_batch_graph_launch(mvs_session *s);
int mvs_batch_graph_check(mvs_session *s);

/* One one-mesh step (mvs_dev_pairvote then mvs_dev_meanshift_votes with these
 * arguments, want_stats off) as a HIP graph owned by the session, with the
 * same record / capture / check rules as mvs_batch_graph_capture: the counts
 * the step reads back (pair total, long sources, mirror total, mean-shift
 * flags) are frozen, so replays are for the captured inputs.  The two
 * mean-shift queries keep their fork onto the side stream inside the graph.
 * mvs_graph_launch / mvs_graph_check replay and check whichever step the
 * session holds (the batch ones are the same calls).  No reference
 * counterpart: the reference runs every stage eagerly
 * (src/core/pipeline.cpp:5-52); a caller that repeats a step over resident
 * inputs (a benchmark, a serving loop) replays it without host
 * synchronisations or per-kernel launches. */
int mvs_step_graph_capture(mvs_session *s, uint64_t n, const double *x, const double *y,
                           const double *z, const double *feat, uint64_t S,
                           const int32_t *selected, double box, const double center[3],
                           double threshold, uint64_t max_pairs, uint64_t max_votes,
                           int precision, uint64_t seeds, uint64_t seed_value, int iterations,
                           double h, double cutoff, mvs_pairvote_result *pv,
                           mvs_meanshift_result *ms);
int mvs_graph_launch(mvs_session *s);
int mvs_graph_check(mvs_session *s);

/* Launch table of a session: pair-walk work quota (sources a persistent wave
 * claims at a time; 0 = one wave per source) and resident blocks, and the
 * seed count below which a mean-shift seed gets a whole 256-lane block.
 * mvs_batch_autotune captures the batch step under each candidate table,
 * times three graph replays each (median, ms), writes every candidate with
 * its time to `table` (capacity >= 16), keeps the fastest on the session and
 * leaves the step captured with it (results in *pv, *ms as for
 * mvs_batch_graph_capture).  Pairs and votes do not depend on the table;
 * mean-shift modes only through the order of their FP sums (16, 32 or 64
 * lanes, or a block, per seed). */
typedef struct {
  uint32_t walk_quota, walk_blocks;
  uint64_t wide_seeds;
  double ms;
  uint32_t seed_lanes;  /* lanes per seed of the narrow queries: 16, 32 or 64 */
  uint32_t reserved;
} mvs_tune_entry;
int mvs_batch_autotune(mvs_session *s, uint64_t n, const double *x, const double *y,
                       const double *z, const double *feat, uint32_t meshes,
                       const uint64_t *sample_offsets, const int32_t *selected,
                       const double *box, const double *center, double threshold,
                       uint64_t max_pairs, uint64_t max_votes, int precision,
                       double seeds_per_sample, uint64_t seed_value, int iterations, double h,
                       double cutoff, mvs_tune_entry *table, uint32_t table_capacity,
                       uint32_t *entries, uint32_t *best, mvs_batch_pairvote_result *pv,
                       mvs_batch_meanshift_result *ms);
int mvs_session_get_tuning(mvs_session *s, mvs_tune_entry *out);
int mvs_session_set_tuning(mvs_session *s, const mvs_tune_entry *in);

/* Mode refinement (no reference counterpart for the energy).
 * mvs_bvh_build keeps a BVH of the mesh surface on the session: median split
 * of face centroids on the widest box axis, at most 8 faces per leaf, as the
 * reference's verification surface (src/cpu/surface.h:12-96); coordinates
 * relative to `center`, in FP32; 1 to 2^27 faces.  Host arrays in.
 * mvs_dev_closest_points: for n centre-relative FP32 points (3 x n SoA) the
 * closest surface point q (3 x n), the unit normal of its face (3 x n), the
 * squared distance and the face id (optional); device arrays.
 * mvs_dev_closest_modes: the same query for the transformed points of many
 * modes in one launch: point j of mode m is mapped by the mode's transform
 * (kind, theta and base as for mvs_dev_align_energy) before the query; pts
 * is modes x 3 x npts, or one 3 x npts set that every mode maps when
 * shared_pts != 0; q and nrm are modes x 3 x npts, dist2 modes x npts.
 * face (optional, modes x npts, in and out): a face id per point that
 * starts the search (its distance bounds the tree walk; -1 for none), then
 * the face of the closest point.  Passing the previous step's faces makes
 * repeated queries of slowly moving points cheap; the result is the same
 * nearest point (between faces at exactly equal distance, the given one).
 * mvs_dev_align_energy: per mode, the point-to-plane energy
 * E = sum_j (n_j . (T(p_j) - q_j))^2 over its npts points (pts, q, nrm:
 * modes x 3 x npts, centre-relative FP32) and its gradient (modes x P) and
 * Gauss-Newton matrix 2 sum J^T J (modes x P x P, optional):
 *   kind 0, rigid, P = 6: theta = (w, v), T(p) = exp([w]x) (R0 p + t0) + v,
 *          base = modes x 12: R0 row-major then t0;
 *   kind 1, mirror, P = 4: theta = (m, d), T(p) = p - 2 (u.p + d) u, u = m/|m|.
 * FP32 arithmetic, sums in a fixed order. */
int mvs_bvh_build(mvs_session *s, uint64_t nv, const double *x, const double *y,
                  const double *z, uint64_t nf, const int32_t *f0, const int32_t *f1,
                  const int32_t *f2, const double center[3]);
int mvs_dev_closest_points(mvs_session *s, uint64_t n, const float *pts, float *q, float *nrm,
                           float *dist2, int32_t *face);
int mvs_dev_closest_modes(mvs_session *s, int kind, uint32_t modes, uint32_t npts,
                          const float *theta, const float *base, const float *pts,
                          int shared_pts, float *q, float *nrm, float *dist2, int32_t *face);
int mvs_dev_align_energy(mvs_session *s, int kind, uint32_t modes, uint32_t npts,
                         const float *theta, const float *base, const float *pts, const float *q,
                         const float *nrm, float *energy, float *grad, float *hess);

/* Batched-seed mean-shift (no reference counterpart; the reference runs
 * sequential seeds only, src/cpu/clustering.cpp:44-60).
 * mvs_dev_ms_points converts compact votes (layout of mvs_dev_vote_emit) to
 * mean-shift points: rigid -> (t_c[3], axis-angle[3]) as a 6*nrigid SoA,
 * reflection -> (d_c, n[3]) as a 4*nrefl SoA, in the record precision.
 * mvs_dev_ms_batched runs `iterations` updates of every seed over one point
 * set (dim 6 or 4, dim*n SoA) with the kernel
 *   w = exp(-d^2/2h^2) - exp(-cutoff^2/2) for d < cutoff*h, else 0,
 * x <- x + sum w (p - x) / sum w.  A seed stops early once an update leaves
 * it unchanged (a fixed point: further updates would repeat it exactly).
 * Seeds and out: dim*ns SoA; weight: ns (sum of weights at the result).
 * MVS_FP32_GUARDED selects FP32 points and accumulation, MVS_FP64 FP64.
 * candidates / in_support / updates (optional) receive the numbers of point
 * visits, of visits inside the support and of updates computed (seeds stop
 * at fixed points, so at most ns * iterations). */
int mvs_dev_ms_points(mvs_session *s, int precision, uint64_t nrigid, const void *q4,
                      const void *t3, uint64_t nrefl, const void *n3, const void *d1,
                      void *rigid_pts, void *refl_pts);
int mvs_dev_ms_batched(mvs_session *s, int precision, int dim, uint64_t n, const void *pts,
                       uint64_t ns, const void *seeds, int iterations, double h, double cutoff,
                       void *out, void *weight, uint64_t *candidates, uint64_t *in_support,
                       uint64_t *updates);

/* mvs_dev_ms_batched that allocates nothing: no hipMalloc, hipFree or pinned
 * host buffer in the call; its scratch is a 16 MiB static device array of
 * the library, one per GPU, kept until the process ends.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  At most
 * MVS_NOALLOC_MAX_POINTS points and MVS_NOALLOC_MAX_SEEDS seeds (larger:
 * MVS_EINVAL); same kernels and results as mvs_dev_ms_batched, without the
 * visit statistics; synchronises once (the non-finite-input check).  Calls
 * that overlap on one GPU share the array and race. */
#define MVS_NOALLOC_MAX_POINTS 65536
#define MVS_NOALLOC_MAX_SEEDS 8192
int mvs_dev_ms_batched_noalloc(mvs_session *s, int precision, int dim, uint64_t n,
                               const void *pts, uint64_t ns, const void *seeds, int iterations,
                               double h, double cutoff, void *out, void *weight);

#ifdef __cplusplus
}
#endif
#endif
