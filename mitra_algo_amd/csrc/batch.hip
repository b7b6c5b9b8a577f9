// Batches of meshes through the voting core (no reference counterpart: the
// reference runs one mesh per detect call, src/core/pipeline.cpp:5-52).
//
// The meshes' vertices and features are concatenated; the selected samples
// of mesh m are ordinals [soff[m], soff[m + 1]).  One pair + vote pass runs
// over every sample of every mesh: the signature sort key carries the mesh
// id above the curvature key, windows stay inside the mesh's key range, the
// walk rejects partners past the mesh's last ordinal, and distances, centres
// and guard bands come from the source's mesh.  Pairs of one mesh are a
// contiguous run in CPU order, and so are its rigid and mirror votes; the
// per-mesh runs equal a one-mesh pass over that mesh.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>

#include "pairing.hpp"

namespace mvs {
namespace {

constexpr int kThreads = 256;

// seg[i] = the mesh whose ordinal range holds i (last m with soff[m] <= i).
__global__ void k_seg_fill(const uint64_t *soff, uint32_t meshes, uint64_t S, uint32_t *seg) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= S)
    return;
  uint32_t f = 0, c = meshes + 1;
  while (c) {
    const uint32_t h = c / 2, m = f + h;
    if (soff[m] <= i) {
      f = m + 1;
      c -= h + 1;
    } else
      c = h;
  }
  seg[i] = f - 1;
}

// Per-mesh starts of the pair list (from the per-source offsets) and of the
// mirror stream (first mirror vote whose pair is in the mesh).
// Offsets are clamped to the pair count P the step was sized with.
__global__ void k_mesh_offsets(const uint64_t *soff, const uint64_t *pair_off, uint32_t meshes,
                               uint64_t P, const int32_t *refl_pair, uint64_t R, uint64_t *poff,
                               uint64_t *roff) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > meshes)
    return;
  const uint64_t p = pair_off[soff[m]] < P ? pair_off[soff[m]] : P;
  poff[m] = p;
  uint64_t f = 0, c = R;
  while (c) {
    const uint64_t h = c / 2, k = f + h;
    if (static_cast<uint64_t>(refl_pair[k]) < p) {
      f = k + 1;
      c -= h + 1;
    } else
      c = h;
  }
  roff[m] = f;
}

// Makes off[0..n) non-decreasing (a running maximum) and at most cap, in
// one block.  Every consumer of per-mesh offsets takes off[m + 1] - off[m]
// as a count; a graph replay over changed inputs searches streams whose
// tails still hold the captured step's entries, so offsets found there need
// not be sorted.  Sorted offsets within capacity pass unchanged.
constexpr int kMonoThreads = 1024;
__global__ void __launch_bounds__(kMonoThreads) k_monotone(uint64_t *off, uint64_t n,
                                                           uint64_t cap) {
  __shared__ uint64_t sh[kMonoThreads];
  uint64_t carry = 0;
  for (uint64_t base = 0; base < n; base += kMonoThreads) {
    const uint64_t i = base + threadIdx.x;
    uint64_t v = i < n ? off[i] : 0;
    v = v < cap ? v : cap;
    v = v > carry ? v : carry;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kMonoThreads; o <<= 1) {
      const uint64_t t = threadIdx.x >= static_cast<unsigned>(o) ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      v = v > t ? v : t;
      sh[threadIdx.x] = v;
      __syncthreads();
    }
    if (i < n)
      off[i] = v;
    carry = sh[kMonoThreads - 1];
    __syncthreads();
  }
}

// Seeds per mesh: ns = min(V, ceil(ratio S)) split between the rigid and
// mirror streams as the one-mesh step splits them (mirror share rounded,
// then each capped by its stream).  Rigid point segments start at 2 x the
// mesh's first pair.
__global__ void k_batch_seeds(const uint64_t *soff, const uint64_t *poff, const uint64_t *roff,
                              uint32_t meshes, double ratio, uint64_t *kr, uint64_t *kf,
                              uint64_t *rseg) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > meshes)
    return;
  rseg[m] = 2 * poff[m];
  if (m == meshes)
    return;
  const uint64_t S = soff[m + 1] - soff[m], nr = 2 * (poff[m + 1] - poff[m]);
  const uint64_t nf = roff[m + 1] - roff[m], V = nr + nf;
  const double want = ceil(ratio * static_cast<double>(S));
  const uint64_t ns = want < static_cast<double>(V) ? static_cast<uint64_t>(want) : V;
  uint64_t f = V ? static_cast<uint64_t>(llround(static_cast<double>(ns) * nf / V)) : 0;
  f = f < nf ? f : nf;
  const uint64_t r = ns > f ? ns - f : 0;
  kr[m] = r < nr ? r : nr;
  kf[m] = f;
}

}  // namespace

// Uploads the batch's per-mesh metadata (offsets, boxes, centres) through
// the session's dedicated pinned buffer and builds the per-ordinal mesh ids.
// Returns the segment view and the smallest box (for the guard bands).
BatchLayout batch_setup(Session *s, uint32_t meshes, const uint64_t *soff, const double *box,
                        const double *center) {
  if (meshes == 0 || !soff || !box || !center)
    raise(MVS_EINVAL, "a batch needs at least one mesh and its offsets, boxes and centres");
  if (soff[0] != 0)
    raise(MVS_EINVAL, "sample offsets must start at 0");
  double min_box = std::numeric_limits<double>::infinity();
  for (uint32_t m = 0; m < meshes; ++m) {
    if (soff[m + 1] < soff[m])
      raise(MVS_EINVAL, "sample offsets must be non-decreasing");
    if (!(box[m] > 0) || !std::isfinite(box[m]))
      raise(MVS_EINVAL, "pairing requires positive finite box size and threshold in (0,1]");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(center[3 * m + k]))
        raise(MVS_EINVAL, "mesh centres must be finite");
    min_box = std::min(min_box, box[m]);
  }
  const uint64_t S = soff[meshes];
  const size_t nb = (meshes + 1) * sizeof(uint64_t) + meshes * sizeof(double) +
                    3 * meshes * sizeof(double);
  BatchLayout L;
  L.meshes = meshes;
  L.S = S;
  L.min_box = min_box;
  char *d = static_cast<char *>(scratch(s, kSlotBatchMeta, nb));
  L.segs.ostart = reinterpret_cast<const uint64_t *>(d);
  L.segs.box = reinterpret_cast<const double *>(d + (meshes + 1) * sizeof(uint64_t));
  L.segs.center = reinterpret_cast<const double *>(d + (meshes + 1) * sizeof(uint64_t) +
                                                   meshes * sizeof(double));
  uint32_t *seg = scratch_as<uint32_t>(s, kSlotBatchSeg, S ? S : 1);
  L.segs.seg = seg;
  s->batch_meshes = meshes;
  s->batch_meta_soff = L.segs.ostart;
  s->batch_soff.assign(soff, soff + meshes + 1);
  s->batch_pv_valid = false;
  if (s->replay == Replay::kCapture)
    return L;  // the recorded step left this batch's metadata and ids in place
  s->drop_graph();  // a captured step would now see other metadata
  // the previous batch's upload may still be reading the pinned buffer
  MVS_HIP(hipStreamSynchronize(s->stream));
  if (nb > s->batch_stage_bytes) {
    if (s->batch_stage)
      (void)hipHostFree(s->batch_stage);
    s->batch_stage = nullptr;
    s->batch_stage_bytes = 0;
    MVS_HIP(hipHostMalloc(&s->batch_stage, nb, hipHostMallocDefault));
    s->batch_stage_bytes = nb;
  }
  char *h = static_cast<char *>(s->batch_stage);
  std::memcpy(h, soff, (meshes + 1) * sizeof(uint64_t));
  std::memcpy(h + (meshes + 1) * sizeof(uint64_t), box, meshes * sizeof(double));
  std::memcpy(h + (meshes + 1) * sizeof(uint64_t) + meshes * sizeof(double), center,
              3 * meshes * sizeof(double));
  MVS_HIP(hipMemcpyAsync(d, h, nb, hipMemcpyHostToDevice, s->stream));
  if (S)
    MVS_LAUNCH(s, "batch_seg", k_seg_fill, blocks_for(S, kThreads), kThreads, 0, L.segs.ostart,
               meshes, S, seg);
  return L;
}

void batch_pairvote(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                    const double *feat, const BatchLayout &L, const int32_t *sel, double t,
                    uint64_t max_pairs, uint64_t max_votes, int precision,
                    mvs_batch_pairvote_result *out) {
  const double no_centre[3] = {0, 0, 0};  // centres come from the segments
  mvs_pairvote_result r{};
  pairvote(s, n, x, y, z, feat, L.S, sel, L.min_box, no_centre, t, max_pairs, max_votes,
           precision, &r, L.segs);
  uint64_t *poff = scratch_as<uint64_t>(s, kSlotBatchPairOff, L.meshes + 1);
  uint64_t *roff = scratch_as<uint64_t>(s, kSlotBatchReflOff, L.meshes + 1);
  MVS_LAUNCH(s, "batch_offsets", k_mesh_offsets, blocks_for(L.meshes + 1, kThreads), kThreads, 0,
             L.segs.ostart, plan_pair_offsets(s), L.meshes, r.pairs, r.refl_pair, r.reflections,
             poff, roff);
  MVS_LAUNCH(s, "batch_offsets", k_monotone, 1, kMonoThreads, 0, poff, L.meshes + 1, r.pairs);
  MVS_LAUNCH(s, "batch_offsets", k_monotone, 1, kMonoThreads, 0, roff, L.meshes + 1,
             r.reflections);
  out->all = r;
  out->meshes = L.meshes;
  out->pair_offsets = poff;
  out->refl_offsets = roff;
  s->batch_pv = *out;
  s->batch_pv_valid = true;
}

void batch_meanshift(Session *s, const uint64_t *soff_host, double ratio, uint64_t seed_value,
                     int iterations, double h, double cutoff, bool want_stats,
                     mvs_batch_meanshift_result *out) {
  if (!s->batch_pv_valid)
    raise(MVS_EINVAL, "batch mean-shift needs a preceding mvs_dev_batch_pairvote");
  if (!(ratio >= 0) || !std::isfinite(ratio))
    raise(MVS_EINVAL, "invalid seed ratio");
  const mvs_batch_pairvote_result &b = s->batch_pv;
  const mvs_pairvote_result &r = b.all;
  const uint32_t meshes = b.meshes;
  const int prec = r.precision;
  const size_t sz = prec == MVS_FP32_GUARDED ? 4 : 8;
  const uint64_t nr = 2 * r.pairs, nf = r.reflections;
  // seed capacity: sum over meshes of ceil(ratio S_m), each stream
  uint64_t cap = 0;
  for (uint32_t m = 0; m < meshes; ++m)
    cap += static_cast<uint64_t>(std::ceil(ratio * static_cast<double>(soff_host[m + 1] -
                                                                       soff_host[m])));
  const uint64_t cr = std::min(cap, nr), cf = std::min(cap, nf);
  uint64_t *kr = scratch_as<uint64_t>(s, kSlotBatchMs4, 6 * (meshes + 1));
  uint64_t *kf = kr + (meshes + 1), *rseg = kf + (meshes + 1);
  uint64_t *offr = rseg + (meshes + 1), *offf = offr + (meshes + 1);
  MVS_LAUNCH(s, "batch_seeds", k_batch_seeds, blocks_for(meshes + 1, kThreads), kThreads, 0,
             s->batch_meta_soff, b.pair_offsets, b.refl_offsets, meshes, ratio, kr, kf, rseg);
  exclusive_scan_u64(s, kr, offr, meshes);
  exclusive_scan_u64(s, kf, offf, meshes);
  // seed slots stay inside the capacities (a replay over changed inputs)
  MVS_LAUNCH(s, "batch_seeds", k_monotone, 1, kMonoThreads, 0, offr, meshes + 1, cr);
  MVS_LAUNCH(s, "batch_seeds", k_monotone, 1, kMonoThreads, 0, offf, meshes + 1, cf);
  int32_t *ids = scratch_as<int32_t>(s, kSlotOutMsSeeds, cr + cf ? cr + cf : 1);
  uint32_t *smesh = scratch_as<uint32_t>(s, kSlotBatchMs5, cr + cf ? cr + cf : 1);
  void *rpts = scratch(s, kSlotOutMsRigidPts, 6 * (nr ? nr : 1) * sz);
  void *fpts = scratch(s, kSlotOutMsReflPts, 4 * (nf ? nf : 1) * sz);
  void *orig = scratch(s, kSlotOutMsRigid, 6 * (cr ? cr : 1) * sz);
  void *owr = scratch(s, kSlotOutMsRigidW, (cr ? cr : 1) * sz);
  void *ofl = scratch(s, kSlotOutMsRefl, 4 * (cf ? cf : 1) * sz);
  void *owf = scratch(s, kSlotOutMsReflW, (cf ? cf : 1) * sz);
  ms_points_device(s, prec, cr ? nr : 0, r.rigid_q4, r.rigid_t3, cf ? nf : 0, r.refl_n3,
                   r.refl_stride, r.refl_d, rpts, fpts);
  MsSegs sr, sf;
  sr.meshes = sf.meshes = meshes;
  sr.seg_off = rseg;
  sr.seed_off = offr;
  sr.seed_mesh = smesh;
  sr.cells_bound = 2 * nr + 1024ULL * meshes;
  sf.seg_off = b.refl_offsets;
  sf.seed_off = offf;
  sf.seed_mesh = smesh + cr;
  sf.cells_bound = 2 * nf + 1024ULL * meshes;
  const uint64_t salt_r = splitmix64(seed_value);
  const uint64_t salt_f = splitmix64(seed_value ^ 0x6d6972726f72ULL);  // "mirror"
  ms_begin(s);
  const int slot_r = want_stats ? 0 : -1, slot_f = want_stats ? 1 : -1;
  if (prec == MVS_FP32_GUARDED) {
    if (cr)
      ms_batched_f32(s, 6, nr, static_cast<const float *>(rpts), cr, nullptr, ids, salt_r,
                     iterations, h, cutoff, static_cast<float *>(orig), static_cast<float *>(owr),
                     slot_r, &sr);
    if (cf)
      ms_batched_f32(s, 4, nf, static_cast<const float *>(fpts), cf, nullptr, ids + cr, salt_f,
                     iterations, h, cutoff, static_cast<float *>(ofl), static_cast<float *>(owf),
                     slot_f, &sf);
  } else {
    if (cr)
      ms_batched_f64(s, 6, nr, static_cast<const double *>(rpts), cr, nullptr, ids, salt_r,
                     iterations, h, cutoff, static_cast<double *>(orig),
                     static_cast<double *>(owr), slot_r, &sr);
    if (cf)
      ms_batched_f64(s, 4, nf, static_cast<const double *>(fpts), cf, nullptr, ids + cr, salt_f,
                     iterations, h, cutoff, static_cast<double *>(ofl),
                     static_cast<double *>(owf), slot_f, &sf);
  }
  unsigned long long sums[kMsSums] = {0, 0, 0, 0, 0, 0};
  ms_finish(s, sums);
  mvs_meanshift_result m{};
  m.rigid_seeds = cr;
  m.refl_seeds = cf;
  m.seed_votes = ids;
  m.rigid_modes = orig;
  m.rigid_weight = owr;
  m.refl_modes = ofl;
  m.refl_weight = owf;
  m.candidates = sums[0] + sums[3];
  m.in_support = sums[1] + sums[4];
  m.updates = sums[2] + sums[5];
  m.rigid_candidates = sums[0];
  m.rigid_in_support = sums[1];
  m.refl_candidates = sums[3];
  m.refl_in_support = sums[4];
  m.precision = prec;
  out->all = m;
  out->meshes = meshes;
  out->rigid_seed_offsets = offr;
  out->refl_seed_offsets = offf;
}

// Everything one batch step needs, as passed to the C entry points.
struct BatchStepArgs {
  uint64_t n;
  const double *x, *y, *z, *feat;
  uint32_t meshes;
  const uint64_t *soff;
  const int32_t *sel;
  const double *box, *center;
  double threshold;
  uint64_t max_pairs, max_votes;
  int precision;
  double seeds_per_sample;
  uint64_t seed_value;
  int iterations;
  double h, cutoff;
};

static void batch_step(Session *s, const BatchStepArgs &a, mvs_batch_pairvote_result *pv,
                       mvs_batch_meanshift_result *ms) {
  BatchLayout L = batch_setup(s, a.meshes, a.soff, a.box, a.center);
  batch_pairvote(s, a.n, a.x, a.y, a.z, a.feat, L, a.sel, a.threshold, a.max_pairs, a.max_votes,
                 a.precision, pv);
  batch_meanshift(s, a.soff, a.seeds_per_sample, a.seed_value, a.iterations, a.h, a.cutoff,
                  false, ms);
}

// Records the step eagerly, then captures it on the session's own stream with
// the recorded counts frozen (see Replay).  `step(false)` is the recorded run
// (its results are the caller's), `step(true)` the captured one.
void graph_capture(Session *s, const std::function<void(bool)> &step) {
  s->drop_graph();
  hipStream_t user = s->stream;
  MVS_HIP(hipStreamSynchronize(user));  // the caller's inputs are ready
  const bool timing = s->timing;
  struct Restore {
    Session *s;
    hipStream_t user;
    bool timing;
    ~Restore() {
      s->stream = user;
      s->replay = Replay::kOff;
      s->timing = timing;
    }
  } restore{s, user, timing};
  s->stream = s->own_stream;  // capture needs a stream other than the legacy one
  s->timing = false;
  s->replay = Replay::kRecord;
  s->rec.clear();
  step(false);
  // recorded values on the device, mismatch word cleared
  s->rec_off.clear();
  size_t total = 0;
  for (const auto &v : s->rec) {
    s->rec_off.push_back(total);
    total += v.size();
  }
  std::vector<unsigned char> flat;
  flat.reserve(total);
  for (const auto &v : s->rec)
    flat.insert(flat.end(), v.begin(), v.end());
  s->rec_dev.ensure(total ? total : 1);
  if (total)
    MVS_HIP(hipMemcpy(s->rec_dev.ptr, flat.data(), total, hipMemcpyHostToDevice));
  s->mismatch_dev = plan_flags(s) + kGraphMismatchWord;
  zero_device(s, s->mismatch_dev, sizeof(unsigned));
  MVS_HIP(hipStreamSynchronize(s->stream));
  s->replay = Replay::kCapture;
  s->rec_pos = 0;
  MVS_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
  hipGraph_t graph = nullptr;
  try {
    step(true);
  } catch (...) {
    (void)hipStreamEndCapture(s->stream, &graph);
    if (graph)
      (void)hipGraphDestroy(graph);
    throw;
  }
  MVS_HIP(hipStreamEndCapture(s->stream, &graph));
  hipError_t e = hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    s->graph_exec = nullptr;
    hip_check(e, "hipGraphInstantiate");
  }
}

static void batch_graph_capture(Session *s, const BatchStepArgs &a, mvs_batch_pairvote_result *pv,
                                mvs_batch_meanshift_result *ms) {
  graph_capture(s, [&](bool captured) {
    if (!captured) {
      batch_step(s, a, pv, ms);
      return;
    }
    mvs_batch_pairvote_result pv2{};
    mvs_batch_meanshift_result ms2{};
    batch_step(s, a, &pv2, &ms2);
  });
}

// Replays the captured step on the session's own stream, ordered after the
// caller's stream and before its later work.
void graph_launch(Session *s) {
  if (!s->graph_exec)
    raise(MVS_EINVAL, "no captured batch step on this session");
  if (!s->graph_events[0]) {
    MVS_HIP(hipEventCreateWithFlags(&s->graph_events[0], hipEventDisableTiming));
    MVS_HIP(hipEventCreateWithFlags(&s->graph_events[1], hipEventDisableTiming));
  }
  const bool other = s->stream != s->own_stream;
  if (other) {
    MVS_HIP(hipEventRecord(s->graph_events[0], s->stream));
    MVS_HIP(hipStreamWaitEvent(s->own_stream, s->graph_events[0], 0));
  }
  MVS_HIP(hipGraphLaunch(s->graph_exec, s->own_stream));
  if (other) {
    MVS_HIP(hipEventRecord(s->graph_events[1], s->own_stream));
    MVS_HIP(hipStreamWaitEvent(s->stream, s->graph_events[1], 0));
  }
}

void graph_check(Session *s) {
  if (!s->graph_exec)
    raise(MVS_EINVAL, "no captured batch step on this session");
  MVS_HIP(hipStreamSynchronize(s->own_stream));
  unsigned mm = 0;
  MVS_HIP(hipMemcpy(&mm, s->mismatch_dev, sizeof mm, hipMemcpyDeviceToHost));
  if (mm) {
    const unsigned zero = 0;
    MVS_HIP(hipMemcpy(s->mismatch_dev, &zero, sizeof zero, hipMemcpyHostToDevice));
    raise(MVS_ERUNTIME,
          "a graph replay read counts other than the captured step's: its inputs changed; "
          "capture the step again");
  }
}

// Autotuner: captures the batch step under each candidate launch table and
// times graph replays (HIP events on the replay stream, median of three);
// the fastest table stays on the session, with its graph captured.
static const Tuning kCandidates[] = {
    // wide seeds (below this many, a mean-shift seed gets a whole block),
    // walk quota, walk blocks (one wave per source, or persistent claims of
    // 1-16 sources from 1024-4096 resident blocks), lanes per narrow seed
    {8192, 0, 2048, 16},  {8192, 0, 2048, 32},  {8192, 0, 2048, 64}, {8192, 1, 2048, 16},
    {8192, 4, 2048, 16},  {8192, 16, 2048, 16}, {8192, 4, 1024, 16}, {8192, 4, 4096, 16},
    {0, 0, 2048, 16},     {0, 4, 2048, 16},     {65536, 0, 2048, 16}, {65536, 4, 2048, 16},
};

static void batch_autotune(Session *s, const BatchStepArgs &a, mvs_tune_entry *table,
                           uint32_t cap, uint32_t *entries, uint32_t *best,
                           mvs_batch_pairvote_result *pv, mvs_batch_meanshift_result *ms) {
  const uint32_t n = sizeof(kCandidates) / sizeof(kCandidates[0]);
  if (!table || cap < n || !entries || !best)
    raise(MVS_EINVAL, "the tuning table needs room for every candidate");
  hipEvent_t ev[2];
  MVS_HIP(hipEventCreate(&ev[0]));
  MVS_HIP(hipEventCreate(&ev[1]));
  struct Events {
    hipEvent_t *e;
    ~Events() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } guard{ev};
  uint32_t arg = 0;
  for (uint32_t c = 0; c < n; ++c) {
    s->tune = kCandidates[c];
    batch_graph_capture(s, a, pv, ms);
    graph_launch(s);  // warm
    graph_check(s);
    float t[3];
    for (int r = 0; r < 3; ++r) {
      MVS_HIP(hipEventRecord(ev[0], s->own_stream));
      MVS_HIP(hipGraphLaunch(s->graph_exec, s->own_stream));
      MVS_HIP(hipEventRecord(ev[1], s->own_stream));
      MVS_HIP(hipEventSynchronize(ev[1]));
      MVS_HIP(hipEventElapsedTime(&t[r], ev[0], ev[1]));
    }
    graph_check(s);
    std::sort(t, t + 3);
    table[c].walk_quota = kCandidates[c].walk_quota;
    table[c].walk_blocks = kCandidates[c].walk_blocks;
    table[c].wide_seeds = kCandidates[c].wide_seeds;
    table[c].seed_lanes = kCandidates[c].seed_lanes;
    table[c].reserved = 0;
    table[c].ms = t[1];
    if (table[c].ms < table[arg].ms)
      arg = c;
  }
  *entries = n;
  *best = arg;
  s->tune = kCandidates[arg];
  batch_graph_capture(s, a, pv, ms);
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_batch_pairvote(mvs_session *sp, uint64_t n, const double *x, const double *y,
                           const double *z, const double *feat, uint32_t meshes,
                           const uint64_t *sample_offsets, const int32_t *selected,
                           const double *box, const double *center, double threshold,
                           uint64_t max_pairs, uint64_t max_votes, int precision,
                           mvs_batch_pairvote_result *out) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!out)
      raise(MVS_EINVAL, "null argument");
    BatchLayout L = batch_setup(s, meshes, sample_offsets, box, center);
    batch_pairvote(s, n, x, y, z, feat, L, selected, threshold, max_pairs, max_votes, precision,
                   out);
  });
}

int mvs_dev_batch_meanshift(mvs_session *sp, double seeds_per_sample, uint64_t seed_value,
                            int iterations, double h, double cutoff, int want_stats,
                            mvs_batch_meanshift_result *out) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (!out)
      raise(MVS_EINVAL, "null argument");
    batch_meanshift(s, s->batch_soff.data(), seeds_per_sample, seed_value, iterations, h, cutoff,
                    want_stats != 0, out);
  });
}

int mvs_batch_graph_capture(mvs_session *sp, uint64_t n, const double *x, const double *y,
                            const double *z, const double *feat, uint32_t meshes,
                            const uint64_t *sample_offsets, const int32_t *selected,
                            const double *box, const double *center, double threshold,
                            uint64_t max_pairs, uint64_t max_votes, int precision,
                            double seeds_per_sample, uint64_t seed_value, int iterations,
                            double h, double cutoff, mvs_batch_pairvote_result *pv,
                            mvs_batch_meanshift_result *ms) {
  return guarded([&] {
    if (!pv || !ms)
      raise(MVS_EINVAL, "null argument");
    BatchStepArgs a{n,         x,        y,         z,         feat,      meshes,
                    sample_offsets,      selected,  box,       center,    threshold,
                    max_pairs, max_votes, precision, seeds_per_sample, seed_value,
                    iterations, h,       cutoff};
    batch_graph_capture(as_session(sp), a, pv, ms);
  });
}

int mvs_batch_autotune(mvs_session *sp, uint64_t n, const double *x, const double *y,
                       const double *z, const double *feat, uint32_t meshes,
                       const uint64_t *sample_offsets, const int32_t *selected,
                       const double *box, const double *center, double threshold,
                       uint64_t max_pairs, uint64_t max_votes, int precision,
                       double seeds_per_sample, uint64_t seed_value, int iterations, double h,
                       double cutoff, mvs_tune_entry *table, uint32_t table_capacity,
                       uint32_t *entries, uint32_t *best, mvs_batch_pairvote_result *pv,
                       mvs_batch_meanshift_result *ms) {
  return guarded([&] {
    if (!pv || !ms)
      raise(MVS_EINVAL, "null argument");
    BatchStepArgs a{n,         x,        y,         z,         feat,      meshes,
                    sample_offsets,      selected,  box,       center,    threshold,
                    max_pairs, max_votes, precision, seeds_per_sample, seed_value,
                    iterations, h,       cutoff};
    batch_autotune(as_session(sp), a, table, table_capacity, entries, best, pv, ms);
  });
}

int mvs_session_get_tuning(mvs_session *sp, mvs_tune_entry *out) {
  return guarded([&] {
    if (!out)
      raise(MVS_EINVAL, "null argument");
    const Tuning &t = as_session(sp)->tune;
    out->walk_quota = t.walk_quota;
    out->walk_blocks = t.walk_blocks;
    out->wide_seeds = t.wide_seeds;
    out->seed_lanes = t.seed_lanes;
    out->reserved = 0;
    out->ms = 0;
  });
}

int mvs_session_set_tuning(mvs_session *sp, const mvs_tune_entry *in) {
  return guarded([&] {
    if (!in || in->walk_blocks == 0 ||
        (in->seed_lanes != 16 && in->seed_lanes != 32 && in->seed_lanes != 64))
      raise(MVS_EINVAL, "invalid tuning entry");
    Session *s = as_session(sp);
    s->drop_graph();  // a captured step keeps the launch table it was captured with
    s->tune.walk_quota = in->walk_quota;
    s->tune.walk_blocks = in->walk_blocks;
    s->tune.wide_seeds = in->wide_seeds;
    s->tune.seed_lanes = in->seed_lanes;
  });
}

int mvs_batch_graph_launch(mvs_session *sp) {
  return guarded([&] { graph_launch(as_session(sp)); });
}

int mvs_batch_graph_check(mvs_session *sp) {
  return guarded([&] { graph_check(as_session(sp)); });
}

}  // extern "C"
