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
__global__ void k_mesh_offsets(const uint64_t *soff, const uint64_t *pair_off, uint32_t meshes,
                               const int32_t *refl_pair, uint64_t R, uint64_t *poff,
                               uint64_t *roff) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > meshes)
    return;
  const uint64_t p = pair_off[soff[m]];
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
  char *d = static_cast<char *>(scratch(s, kSlotBatchMeta, nb));
  MVS_HIP(hipMemcpyAsync(d, h, nb, hipMemcpyHostToDevice, s->stream));
  BatchLayout L;
  L.meshes = meshes;
  L.S = S;
  L.min_box = min_box;
  L.segs.ostart = reinterpret_cast<const uint64_t *>(d);
  L.segs.box = reinterpret_cast<const double *>(d + (meshes + 1) * sizeof(uint64_t));
  L.segs.center = reinterpret_cast<const double *>(d + (meshes + 1) * sizeof(uint64_t) +
                                                   meshes * sizeof(double));
  uint32_t *seg = scratch_as<uint32_t>(s, kSlotBatchSeg, S ? S : 1);
  if (S)
    MVS_LAUNCH(s, "batch_seg", k_seg_fill, blocks_for(S, kThreads), kThreads, 0, L.segs.ostart,
               meshes, S, seg);
  L.segs.seg = seg;
  s->batch_meshes = meshes;
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
             L.segs.ostart, plan_pair_offsets(s), L.meshes, r.refl_pair, r.reflections, poff,
             roff);
  out->all = r;
  out->meshes = L.meshes;
  out->pair_offsets = poff;
  out->refl_offsets = roff;
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

}  // extern "C"
