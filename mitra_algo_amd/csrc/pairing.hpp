// Shared between the pairing and vote translation units.
#pragma once
#include "common.hpp"
#include "session.hpp"

namespace mvs {

// Ordinal-order sample SoA lanes (14 per sample): position, kmin, kmax,
// normal, dmin, dmax.  The FP32 copy holds positions relative to the model
// centre.
enum SampleLane {
  L_X = 0, L_Y, L_Z, L_K0, L_K1, L_NX, L_NY, L_NZ, L_DX, L_DY, L_DZ, L_XX, L_XY, L_XZ, L_N
};

// Layout of the small flags slot shared by the pairing and vote phases
// (kFlagsBytes bytes): 32-bit words 0 error flags, 1 long-source count,
// 2-3 pair total (uint64), kRmaxWord the FP32 coordinate bound (float bits);
// 64-bit word kVoteTotalWord the mirror-vote total.
constexpr size_t kFlagsBytes = 256;
constexpr int kRmaxWord = 8;
constexpr int kVoteTotalWord = 8;
inline unsigned *plan_flags(Session *s) {
  return static_cast<unsigned *>(scratch(s, kSlotFlags, kFlagsBytes));
}

// Synchronous read-back of a few device words through pinned staging.
void read_back(Session *s, void *dst, const void *src, size_t bytes);

// Pairing phases (pairing.hip).
void pair_count_device(Session *s, uint64_t n, const double *x, const double *y, const double *z,
                       const double *feat, uint64_t S, const int32_t *sel, double box, double t,
                       uint64_t max_pairs, int precision, const double *center, uint64_t *P_out);
void pair_fill_device(Session *s, int32_t *a, int32_t *b, bool ordinals);
// Sums of per-source statistics of the last count: window candidates,
// 2-D index candidates and FP64 rechecks.
void pair_statistics(Session *s, uint64_t out[3]);

// Sample SoA of the last pair count (valid until the next count).
const double *plan_samples64(Session *s);
const float *plan_samples32(Session *s);

}  // namespace mvs
