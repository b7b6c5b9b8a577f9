// Fused device entry points of the voting step: pairing + vote emission in
// one call, then batched mean-shift over the emitted votes.  Outputs live in
// session-owned buffers, so a caller's step is two calls with two short
// host synchronisations (the pair total for the budget check, the mirror
// total at the end).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "pairing.hpp"

namespace mvs {

void pairvote(Session *s, uint64_t n, const double *x, const double *y, const double *z,
              const double *feat, uint64_t S, const int32_t *sel, double box,
              const double *center, double t, uint64_t max_pairs, uint64_t max_votes,
              int precision, mvs_pairvote_result *out, const Segs &segs) {
  s->pairvote_valid = false;
  s->batch_pv_valid = false;  // a batch result's buffers are about to be reused
  uint64_t P = 0;
  pair_count_device(s, n, x, y, z, feat, S, sel, box, t, max_pairs, precision, center, &P,
                    segs);
  if (2 * P > max_votes)
    raise(MVS_EBUDGET, "Vote budget exceeded; lower --sample-ratio or tighten --pair-threshold");
  const size_t sz = precision == MVS_FP32_GUARDED ? 4 : 8;
  const uint64_t Pp = P ? P : 1;
  int32_t *ab = scratch_as<int32_t>(s, kSlotOutPairs, 2 * Pp);
  void *q4 = scratch(s, kSlotOutRigidQ, 4 * 2 * Pp * sz);
  void *t3 = scratch(s, kSlotOutRigidT, 3 * 2 * Pp * sz);
  void *n3 = scratch(s, kSlotOutReflN, 3 * Pp * sz);
  void *d1 = scratch(s, kSlotOutReflD, Pp * sz);
  int32_t *fp = scratch_as<int32_t>(s, kSlotOutReflPair, Pp);
  pair_fill_device(s, ab, ab + Pp, true);
  uint64_t *rtot = reinterpret_cast<uint64_t *>(plan_flags(s)) + kVoteTotalWord;
  vote_offsets_device(s, P, ab, ab + Pp, precision, rtot, segs, true);
  // no rigid pair-id stream: rigid vote v belongs to pair v / 2
  vote_emit_strided(s, P, ab, ab + Pp, center, precision, Pp, q4, t3, nullptr, n3, d1, fp,
                    segs);
  const uint64_t R = vote_total_checked(s);
  if (2 * P + R > max_votes || 2 * P + R > static_cast<uint64_t>(INT32_MAX))
    raise(MVS_EBUDGET, "Vote budget exceeded; lower --sample-ratio or tighten --pair-threshold");
  mvs_pairvote_result r{};
  r.pairs = P;
  r.reflections = R;
  r.a = ab;
  r.b = ab + Pp;
  r.rigid_q4 = q4;
  r.rigid_t3 = t3;
  r.rigid_pair = nullptr;
  r.refl_n3 = n3;
  r.refl_d = d1;
  r.refl_pair = fp;
  r.refl_stride = Pp;
  r.precision = precision;
  s->pairvote = r;
  s->pairvote_valid = true;
  s->vote_plan.P = P;
  s->vote_plan.R = R;
  s->vote_plan.precision = precision;
  s->vote_plan.valid = true;
  *out = r;
}

static void meanshift_votes(Session *s, uint64_t nseeds, uint64_t seed_value, int iterations,
                            double h, double cutoff, bool want_stats, mvs_meanshift_result *out) {
  if (!s->pairvote_valid)
    raise(MVS_EINVAL, "mean-shift over votes needs a preceding mvs_dev_pairvote");
  const mvs_pairvote_result &r = s->pairvote;
  const uint64_t nr = 2 * r.pairs, nf = r.reflections, V = nr + nf;
  const int prec = r.precision;
  const size_t sz = prec == MVS_FP32_GUARDED ? 4 : 8;
  uint64_t kf = V ? static_cast<uint64_t>(std::llround(static_cast<double>(nseeds) * nf / V)) : 0;
  kf = std::min(kf, nf);
  const uint64_t kr = std::min(nseeds > kf ? nseeds - kf : 0, nr);
  int32_t *dseeds = scratch_as<int32_t>(s, kSlotOutMsSeeds, kr + kf ? kr + kf : 1);
  const uint64_t salt_r = splitmix64(seed_value);
  const uint64_t salt_f = splitmix64(seed_value ^ 0x6d6972726f72ULL);  // "mirror"
  void *rpts = scratch(s, kSlotOutMsRigidPts, 6 * (nr ? nr : 1) * sz);
  void *fpts = scratch(s, kSlotOutMsReflPts, 4 * (nf ? nf : 1) * sz);
  void *orig = scratch(s, kSlotOutMsRigid, 6 * (kr ? kr : 1) * sz);
  void *owr = scratch(s, kSlotOutMsRigidW, (kr ? kr : 1) * sz);
  void *ofl = scratch(s, kSlotOutMsRefl, 4 * (kf ? kf : 1) * sz);
  void *owf = scratch(s, kSlotOutMsReflW, (kf ? kf : 1) * sz);
  if (prec != MVS_FP32_GUARDED && prec != MVS_FP64)
    raise(MVS_EINVAL, "unknown precision");
  // the rigid kind whole on the side stream (its own scratch bank), queued
  // first (measured: faster than the mirror kind first), beside the mirror
  // kind on the session stream: the two set-ups run together; each query is
  // held until both set-ups are queued (ms_release), the mirror one, the
  // longer, then starts as soon as its set-up is done and the rigid one
  // behind both set-ups.  A set-up queued behind a running query waits for
  // its blocks to drain and is slowed several times over.  Each run writes
  // its kind's points from the votes in the launch that takes their box and
  // moments.
  ms_begin(s, kr && kf);
  MsRun rigid_run = kf ? MsRun{1, true} : MsRun{};
  rigid_run.hold = true;
  rigid_run.gate = true;
  rigid_run.from_a = r.rigid_q4, rigid_run.from_b = r.rigid_t3;
  MsRun mirror_run{};
  mirror_run.hold = true;
  mirror_run.from_a = r.refl_n3, mirror_run.from_b = r.refl_d;
  mirror_run.from_stride = r.refl_stride;
  const int slot_r = want_stats ? 0 : -1, slot_f = want_stats ? 1 : -1;
  auto rigid = [&] {
    if (!kr)
      return;
    if (prec == MVS_FP32_GUARDED)
      ms_batched_f32(s, 6, nr, static_cast<const float *>(rpts), kr, nullptr, dseeds, salt_r,
                     iterations, h, cutoff, static_cast<float *>(orig), static_cast<float *>(owr),
                     slot_r, nullptr, rigid_run);
    else
      ms_batched_f64(s, 6, nr, static_cast<const double *>(rpts), kr, nullptr, dseeds, salt_r,
                     iterations, h, cutoff, static_cast<double *>(orig),
                     static_cast<double *>(owr), slot_r, nullptr, rigid_run);
  };
  auto mirror = [&] {
    if (!kf)
      return;
    if (prec == MVS_FP32_GUARDED)
      ms_batched_f32(s, 4, nf, static_cast<const float *>(fpts), kf, nullptr, dseeds + kr,
                     salt_f, iterations, h, cutoff, static_cast<float *>(ofl),
                     static_cast<float *>(owf), slot_f, nullptr, mirror_run);
    else
      ms_batched_f64(s, 4, nf, static_cast<const double *>(fpts), kf, nullptr, dseeds + kr,
                     salt_f, iterations, h, cutoff, static_cast<double *>(ofl),
                     static_cast<double *>(owf), slot_f, nullptr, mirror_run);
  };
  rigid();
  mirror();
  unsigned long long sums[kMsSums] = {0, 0, 0, 0, 0, 0};
  ms_finish(s, sums);
  mvs_meanshift_result m{};
  m.rigid_seeds = kr;
  m.refl_seeds = kf;
  m.seed_votes = dseeds;
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
  *out = m;
}

// One step of the voting core (pairs + votes, then mean-shift over them).
struct StepArgs {
  uint64_t n;
  const double *x, *y, *z, *feat;
  uint64_t S;
  const int32_t *sel;
  double box;
  double center[3];
  double threshold;
  uint64_t max_pairs, max_votes;
  int precision;
  uint64_t seeds, seed_value;
  int iterations;
  double h, cutoff;
};

static void step_capture(Session *s, const StepArgs &a, mvs_pairvote_result *pv,
                         mvs_meanshift_result *ms) {
  graph_capture(s, [&](bool captured) {
    mvs_pairvote_result pv2{};
    mvs_meanshift_result ms2{};
    pairvote(s, a.n, a.x, a.y, a.z, a.feat, a.S, a.sel, a.box, a.center, a.threshold, a.max_pairs,
             a.max_votes, a.precision, captured ? &pv2 : pv, Segs{});
    meanshift_votes(s, a.seeds, a.seed_value, a.iterations, a.h, a.cutoff, false,
                    captured ? &ms2 : ms);
  });
}

}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_pairvote(mvs_session *sp, uint64_t n, const double *x, const double *y,
                     const double *z, const double *feat, uint64_t S, const int32_t *sel,
                     double box, const double center[3], double t, uint64_t max_pairs,
                     uint64_t max_votes, int precision, mvs_pairvote_result *out) {
  return guarded([&] {
    if (!out || !center)
      raise(MVS_EINVAL, "null argument");
    pairvote(as_session(sp), n, x, y, z, feat, S, sel, box, center, t, max_pairs, max_votes,
             precision, out, Segs{});
  });
}

int mvs_dev_meanshift_votes(mvs_session *sp, uint64_t seeds, uint64_t seed_value, int iterations,
                            double h, double cutoff, int want_stats, mvs_meanshift_result *out) {
  return guarded([&] {
    if (!out)
      raise(MVS_EINVAL, "null argument");
    meanshift_votes(as_session(sp), seeds, seed_value, iterations, h, cutoff, want_stats != 0,
                    out);
  });
}

int mvs_step_graph_capture(mvs_session *sp, uint64_t n, const double *x, const double *y,
                           const double *z, const double *feat, uint64_t S, const int32_t *sel,
                           double box, const double center[3], double t, uint64_t max_pairs,
                           uint64_t max_votes, int precision, uint64_t seeds,
                           uint64_t seed_value, int iterations, double h, double cutoff,
                           mvs_pairvote_result *pv, mvs_meanshift_result *ms) {
  return guarded([&] {
    if (!pv || !ms || !center)
      raise(MVS_EINVAL, "null argument");
    StepArgs a{n, x, y, z, feat, S, sel, box, {center[0], center[1], center[2]}, t, max_pairs,
               max_votes, precision, seeds, seed_value, iterations, h, cutoff};
    step_capture(as_session(sp), a, pv, ms);
  });
}

int mvs_graph_launch(mvs_session *sp) {
  return guarded([&] { graph_launch(as_session(sp)); });
}

int mvs_graph_check(mvs_session *sp) {
  return guarded([&] { graph_check(as_session(sp)); });
}

}  // extern "C"
