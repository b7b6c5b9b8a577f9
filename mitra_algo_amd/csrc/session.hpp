// Session state shared by the stage translation units.
#pragma once
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace mvs {

struct PendingTiming {
  const char *name;
  hipEvent_t start, stop;
};

// Batch layout (several meshes in one call).  Ordinals of mesh m are the
// contiguous range [ostart[m], ostart[m + 1]); vertex ids in the selected
// list index the concatenated vertex and feature arrays.  seg == nullptr
// means one mesh whose box and centre are passed as plain values.
struct Segs {
  const uint32_t *seg = nullptr;     // per ordinal: mesh id
  const uint64_t *ostart = nullptr;  // per mesh + 1: first ordinal (device)
  const double *box = nullptr;       // per mesh: box size (device)
  const double *center = nullptr;    // per mesh: 3 doubles (device)
};

// Device-resident plan kept between the count and fill phases of pairing.
struct PairPlan {
  uint64_t S = 0;          // selected samples
  uint64_t P = 0;          // pairs
  uint64_t long_sources = 0;
  double threshold = 0, box = 0;
  int precision = MVS_FP64;
  const int32_t *selected = nullptr;  // caller's device list (mesh ids)
  double min_box = 0;                 // smallest box of the batch (guard bands)
  Segs segs;
  bool valid = false;
};

struct VotePlan {
  uint64_t P = 0, R = 0;  // pairs and reflection votes
  int precision = MVS_FP64;
  bool valid = false;
};

// Record / replay of a step for graph capture: kRecord runs eagerly and keeps
// every value read back; kCapture runs under stream capture and takes those
// values instead of reading (each site leaves a device check).
enum class Replay { kOff, kRecord, kCapture };

// Launch parameters picked by the autotuner (defaults: the hand-set values).
struct Tuning {
  uint64_t wide_seeds = 8192;  // below this many seeds, a seed gets four waves
  unsigned walk_quota = 0;     // sources per claim of a persistent pair walk (0: one per wave)
  unsigned walk_blocks = 2048; // resident blocks of the persistent pair walk
  unsigned seed_lanes = 32;    // lanes per seed of the narrow mean-shift query (16, 32, 64)
};

// Caller-provided device memory that serves every scratch request of one
// call instead of the session's slots (the allocation-free entry point): each
// slot gets a 256-byte aligned region on its first request in the call, a
// larger later request a new one; nothing is allocated or freed.
struct Arena {
  unsigned char *base = nullptr;
  size_t cap = 0, used = 0;
  size_t off[kSlotCount] = {}, len[kSlotCount] = {};
};

struct Session {
  int device = 0;
  unsigned walk_resident = 0;  // blocks of the static pair walk the device holds at once
  Arena *arena = nullptr;  // set for the duration of an allocation-free call
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  // second stream for independent work forked from `stream` and joined back
  // with events before anything reads its results
  hipStream_t side_stream = nullptr;
  hipEvent_t fork_event = nullptr, join_event = nullptr;
  bool side_pending = false;
  // mean-shift queries held back until every run of the call has queued its
  // set-up (ms_release launches them, each behind the other runs' set-ups)
  struct MsHeld {
    hipStream_t stream;
    int bank;
    bool gate;  // also waits for the other held runs' set-ups
    std::function<void()> launch;
  };
  std::vector<MsHeld> ms_held;
  hipEvent_t setup_events[2] = {nullptr, nullptr};
  // scratch bank of the sort and scan primitives (1: a side-stream run's own
  // slots, so it can overlap bank-0 work)
  int bank = 0;
  DBuf<unsigned char> slots[kSlotCount];
  void *stage = nullptr;
  size_t stage_bytes = 0;
  // pinned copy of the current batch's per-mesh metadata (a captured graph
  // re-reads it at every replay, so nothing else writes it)
  void *batch_stage = nullptr;
  size_t batch_stage_bytes = 0;
  bool timing = false;
  std::string timing_only;  // ",name,name," or empty for every kernel
  bool timing_open = false;  // begin() recorded a start event for the current launch
  std::vector<PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, uint64_t>> timings;
  PairPlan pair_plan;
  VotePlan vote_plan;
  uint64_t batch_meshes = 0;  // mesh count of the current batch call (sort key width)
  const uint64_t *batch_meta_soff = nullptr;  // device copy of the sample offsets
  std::vector<uint64_t> batch_soff;           // host copy
  mvs_batch_pairvote_result batch_pv{};
  bool batch_pv_valid = false;
  Tuning tune;
  // graph of one batch step (see Replay)
  Replay replay = Replay::kOff;
  std::vector<std::vector<unsigned char>> rec;  // values read back while recording
  std::vector<size_t> rec_off;                  // their offsets in rec_dev
  size_t rec_pos = 0;
  DBuf<unsigned char> rec_dev;                  // the recorded values on the device
  unsigned *mismatch_dev = nullptr;             // set by a replay that read other values
  hipGraphExec_t graph_exec = nullptr;
  hipEvent_t graph_events[2] = {nullptr, nullptr};  // orders replays with the caller's stream
  void drop_graph();
  // mesh surface for closest-point queries (centre-relative FP32 BVH)
  DBuf<unsigned char> bvh_nodes, bvh_tri, bvh_fid, bvh_leaf_of;
  uint64_t bvh_faces = 0;
  int bvh_root = 0, bvh_stack = 1;  // root reference; LDS stack entries per thread
  DBuf<float> align_part;            // energy partials: modes x chunks x terms
  // outputs of the last fused pair + vote call (session-owned buffers)
  mvs_pairvote_result pairvote{};
  bool pairvote_valid = false;
  // host-layer results awaiting fetch
  DBuf<int32_t> out_a, out_b, out_pair;
  DBuf<uint8_t> out_kind;
  DBuf<double> out_lanes;
  uint64_t out_count = 0;
  // clustering results awaiting fetch
  std::vector<uint8_t> cl_refl;
  std::vector<double> cl_rot, cl_trans, cl_pn, cl_po;
  std::vector<uint64_t> cl_off;
  std::vector<int32_t> cl_members;
  // verification results awaiting fetch (index 0: side A, 1: side B)
  std::vector<uint64_t> ver_side_off[2], ver_groups[2], ver_patch_off[2];
  std::vector<int32_t> ver_side[2], ver_members[2];
  std::vector<double> ver_cov, ver_rms;
  uint64_t ver_rows = 0;  // per-row results of the last call stay on the device
  uint64_t ver_us[3] = {0, 0, 0};  // host clock: device phase, download, assembly
  std::vector<std::vector<int32_t>> ver_slot;  // per-thread patch scratch
  // verification surface: search records of the reference-order FP64 tree,
  // leaf-ordered faces and their normals
  DBuf<unsigned char> ver_nodes, ver_faces, ver_normals;
  int ver_root = 0;           // root reference (inner node 0, or a leaf)
  double ver_box[6] = {0, 0, 0, 0, 0, 0};  // surface box: lo xyz, hi xyz
  int ver_stack = 0;          // LDS stack entries per thread (tree height + 1)
  uint64_t ver_surface_nv = 0;
  bool ver_surface = false;

  void begin(const char *name);
  void end();
  void collect();
  void check_launch(const char *what);
  ~Session();
};

// Launch helper: records per-kernel HIP events when timing is on.
#define MVS_LAUNCH(sess, name, kernel, grid, block, lds, ...)                    \
  do {                                                                          \
    (sess)->begin(name);                                                         \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, (sess)->stream,     \
                       __VA_ARGS__);                                             \
    (sess)->check_launch(name);                                                  \
    (sess)->end();                                                               \
  } while (0)

}  // namespace mvs
