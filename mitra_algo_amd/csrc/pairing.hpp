// Shared between the pairing and vote translation units.
#pragma once
#include <functional>

#include "common.hpp"
#include "session.hpp"

namespace mvs {

// Ordinal-order sample records (AoS, kRec scalars per sample, 14 used):
// position, kmin, kmax, normal, dmin, dmax, 2 pad.  One FP32 record is one
// 64-byte cache line, an FP64 record two; the FP32 copy holds positions
// relative to the model centre.
enum SampleField {
  L_X = 0, L_Y, L_Z, L_K0, L_K1, L_NX, L_NY, L_NZ, L_DX, L_DY, L_DZ, L_XX, L_XY, L_XZ, L_N
};
constexpr int kRec = 16;

// Layout of the small flags slot shared by the pairing and vote phases
// (kFlagsBytes bytes): 32-bit words 0 error flags, 1 long-source count,
// 2-3 pair total (uint64), kRmaxWord the FP32 coordinate bound (float bits);
// 64-bit word kVoteTotalWord the mirror-vote total, 32-bit word kVoteBadWord
// (inside 64-bit word kVoteTotalWord + 2) the out-of-range ordinal flag,
// 32-bit word kGraphMismatchWord the replay check of a captured step.
constexpr size_t kFlagsBytes = 256;
constexpr int kRmaxWord = 8;
constexpr int kVoteTotalWord = 8;
constexpr int kVoteBadWord = 2 * (kVoteTotalWord + 2);
constexpr int kGraphMismatchWord = 40;  // set by a graph replay that read other values
inline unsigned *plan_flags(Session *s) {
  return static_cast<unsigned *>(scratch(s, kSlotFlags, kFlagsBytes));
}

// Synchronous read-back of a few device words through pinned staging.
void read_back(Session *s, void *dst, const void *src, size_t bytes);

// Pairing phases (pairing.hip).
void pair_count_device(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                       const double *feat, uint64_t S, const int32_t *sel, double box, double t,
                       uint64_t max_pairs, int precision, const double *center, uint64_t *P_out,
                       const Segs &segs = Segs{});
void pair_fill_device(Session *s, int32_t *a, int32_t *b, bool ordinals);
// Sums of per-source statistics of the last count: window candidates,
// 2-D index candidates and FP64 rechecks.
void pair_statistics(Session *s, uint64_t out[3]);

// Vote phases (votes.hip).
void vote_count_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b, int precision,
                       uint64_t *R_out);
void vote_emit_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                      const double *center, int precision, void *q4, void *t3, int32_t *rp,
                      void *n3, void *d1, int32_t *fp);
// Vote phases without the reflection read-back: mirror votes go to a stream
// of capacity (and SoA stride) `stride`; the total is left on the device at
// *total_dev (uint64).  bad_cleared: the out-of-range flag word was cleared
// with the pair-count flags and is not cleared again.
void vote_offsets_device(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                         int precision, uint64_t *total_dev, const Segs &segs = Segs{},
                         bool bad_cleared = false);
void vote_emit_strided(Session *s, uint64_t P, const int32_t *a, const int32_t *b,
                       const double *center, int precision, uint64_t stride, void *q4, void *t3,
                       int32_t *rp, void *n3, void *d1, int32_t *fp, const Segs &segs = Segs{});
// Synchronous read of the mirror total left by vote_offsets_device; raises
// MVS_EINVAL when the count pass met a pair ordinal outside [0, S).
uint64_t vote_total_checked(Session *s);

// Batched mean-shift (msbatch.hip).  A call is ms_begin, up to two runs
// (run_slot k = 0 or 1 receives the sums of candidate visits, in-support
// visits and updates computed at sums[3k..3k+2]; -1 for none), then
// ms_finish: one synchronisation that raises MVS_EINVAL for non-finite
// points and returns the six sums.
constexpr int kMsSums = 6;
// Segments of one mean-shift run (batches of meshes).
struct MsSegs {
  uint32_t meshes = 1;
  const uint64_t *seg_off = nullptr;   // device, meshes + 1: each mesh's points
  const uint64_t *seed_off = nullptr;  // device, meshes + 1: each mesh's seeds
  uint32_t *seed_mesh = nullptr;       // device, seed capacity: mesh of each seed
  uint64_t cells_bound = 0;            // bound on the total grid cells
};
// fork: record the side stream's fork point (needed by a side run).
void ms_begin(Session *s, bool fork = false);
// Launches the held queries (MsRun::hold) of the runs queued since
// ms_begin, each on its run's stream behind the set-ups of the other runs
// (a query that occupies the chip would slow a set-up queued behind it
// several times over); ms_finish calls it too.
void ms_release(Session *s);
void ms_finish(Session *s, unsigned long long sums[kMsSums]);
// Seeds: explicit positions (D x ns SoA), or, with seed_ids, ns of the n
// points drawn stratified over the cell-sorted order with `salt` (indices
// written to seed_ids, in cell order).
// Two runs of one call may overlap: one with set = 1 and side = true runs
// whole (set-up and query) on the session's side stream, forked at
// ms_begin(s, true), in its own slot set and scratch bank; the other runs on
// the session stream with set 0; ms_finish joins the side stream before
// reading anything back.
struct MsRun {
  int set = 0;        // slot set and scratch bank (0 or 1)
  bool side = false;  // the run on the side stream
  bool hold = false;  // the query waits for ms_release
  bool gate = false;  // held: the query also waits for the other held runs' set-ups
  // optional: the votes the points come from (rigid: q4, t3; mirror: n3 with
  // row stride `from_stride`, d1); the run then writes the points into its
  // `pts` argument, in the launch that takes their box and moments
  const void *from_a = nullptr, *from_b = nullptr;
  uint64_t from_stride = 0;
};
void ms_batched_f32(Session *s, int D, uint64_t n, const float *pts, uint64_t ns,
                    const float *seeds, int32_t *seed_ids, uint64_t salt, int iterations,
                    double h, double cutoff, float *out, float *weight, int run_slot,
                    const MsSegs *seg = nullptr, MsRun run = {});
void ms_batched_f64(Session *s, int D, uint64_t n, const double *pts, uint64_t ns,
                    const double *seeds, int32_t *seed_ids, uint64_t salt, int iterations,
                    double h, double cutoff, double *out, double *weight, int run_slot,
                    const MsSegs *seg = nullptr, MsRun run = {});
// The non-finite-input flag of the current mean-shift call (device word).
unsigned *ms_bad_flag(Session *s);
// Compact votes to mean-shift points; mirror normals are 3 rows of n3_stride.
void ms_points_device(Session *s, int precision, uint64_t nrigid, const void *q4, const void *t3,
                      uint64_t nrefl, const void *n3, uint64_t n3_stride, const void *d1,
                      void *rigid_pts, void *refl_pts);

// Fused pair + vote step (fused.hip); segs describes a batch of meshes.
void pairvote(Session *s, uint64_t n, const double *x, const double *y, const double *z,
              const double *feat, uint64_t S, const int32_t *sel, double box,
              const double *center, double t, uint64_t max_pairs, uint64_t max_votes,
              int precision, mvs_pairvote_result *out, const Segs &segs);
// Per-ordinal pair offsets of the last pair count ([S + 1], device).
const uint64_t *plan_pair_offsets(Session *s);

// Batches (batch.hip): per-mesh metadata on the device and the segment view.
struct BatchLayout {
  Segs segs;
  uint32_t meshes = 0;
  uint64_t S = 0;       // samples of all meshes
  double min_box = 0;   // smallest box (guard bands)
};
BatchLayout batch_setup(Session *s, uint32_t meshes, const uint64_t *soff, const double *box,
                        const double *center);
void batch_pairvote(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                    const double *feat, const BatchLayout &L, const int32_t *sel, double t,
                    uint64_t max_pairs, uint64_t max_votes, int precision,
                    mvs_batch_pairvote_result *out);
void batch_meanshift(Session *s, const uint64_t *soff_host, double ratio, uint64_t seed_value,
                     int iterations, double h, double cutoff, bool want_stats,
                     mvs_batch_meanshift_result *out);

// Graph of one step (batch.hip).  graph_capture runs step(false) eagerly,
// recording every value read back, then captures step(true) on the session's
// own stream with those values frozen (each read-back site becomes a device
// check).  graph_launch replays it ordered after the session stream's queued
// work; graph_check synchronises and raises MVS_ERUNTIME when a replay read
// other values.
void graph_capture(Session *s, const std::function<void(bool)> &step);
void graph_launch(Session *s);
void graph_check(Session *s);

// Sample records of the last pair count (valid until the next count).
const double *plan_samples64(Session *s);
const float *plan_samples32(Session *s);

}  // namespace mvs
