// Shared host/device infrastructure of libmvs: status codes, error transport,
// device buffers, the session and small device math helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mvs.h"

namespace mvs {

// ------------------------------------------------------------------ errors
// Exceptions inside the library carry an mvs status code; every extern "C"
// entry point converts them back to (code, thread-local message).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void raise(int code, const std::string &msg) { throw Error(code, msg); }
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    raise(MVS_EHIP, std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}
#define MVS_HIP(call) ::mvs::hip_check((call), #call)
void set_error(int code, const char *msg);
template <class F> int guarded(F &&f) {
  try {
    f();
    return MVS_OK;
  } catch (const Error &e) {
    set_error(e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error(MVS_EHIP, "host allocation failed");
    return MVS_EHIP;
  } catch (const std::exception &e) {
    set_error(MVS_ERUNTIME, e.what());
    return MVS_ERUNTIME;
  }
}

// ---------------------------------------------------------------- buffers
// Device buffer owned by the caller's scope; freed on unwind.
// Device bytes the calling thread's libmvs buffers hold, and their peak
// (session.hip; read through mvs_thread_metrics).
void note_device_alloc(size_t bytes);
void note_device_free(size_t bytes);

template <class T> struct DBuf {
  T *ptr = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  void alloc(size_t count) {
    release();
    if (count > SIZE_MAX / sizeof(T))
      raise(MVS_EOVERFLOW, "device allocation size overflow");
    if (count) {
      MVS_HIP(hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T)));
      note_device_alloc(count * sizeof(T));
    }
    n = count;
  }
  void ensure(size_t count) {
    if (count > n)
      alloc(count);
  }
  void release() {
    if (ptr) {
      (void)hipFree(ptr);
      note_device_free(n * sizeof(T));
    }
    ptr = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr, o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    release();
    ptr = o.ptr, n = o.n;
    o.ptr = nullptr, o.n = 0;
    return *this;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------- session
struct Session;  // defined in session.hip
Session *as_session(mvs_session *s);

// Stream-ordered helpers bound to a session.
hipStream_t session_stream(Session *s);
void *scratch(Session *s, int slot, size_t bytes);  // grow-only per-slot arena
template <class T> T *scratch_as(Session *s, int slot, size_t count) {
  return static_cast<T *>(scratch(s, slot, count * sizeof(T)));
}
void *host_stage(Session *s, size_t bytes);  // pinned staging for small D2H
void note_copy(bool to_device, size_t bytes);
void copy_h2d(Session *s, void *dst, const void *src, size_t bytes);
void copy_d2h(Session *s, void *dst, const void *src, size_t bytes);
void sync(Session *s);

// Scratch slots (each slot is an independent grow-only allocation).
enum Slot : int {
  kSlotSort0 = 0, kSlotSort1, kSlotSort2, kSlotSort3, kSlotSortHist,
  kSlotScan, kSlotFlags, kSlotSamples, kSlotWindows, kSlotCounts, kSlotOffsets,
  kSlotRefl, kSlotBlock, kSlotMisc0, kSlotMisc1, kSlotMisc2, kSlotMisc3,
  // pairing plan (persist between count, fill and vote emission)
  kSlotPlanSamples32, kSlotPlanKeys, kSlotPlanOrd, kSlotPlanSortedK, kSlotPlanSortedK32,
  kSlotPlanWin, kSlotPlanCounts, kSlotPlanOffsets, kSlotPlanLong, kSlotPlanStats,
  kSlotPlanSlab, kSlotPlanKw, kSlotPlanLongCnt, kSlotPlanLongOff,
  kSlotVoteBlock, kSlotVoteBlockOff,
  kSlotMs0, kSlotMs1, kSlotMs2, kSlotMs3, kSlotMs4, kSlotMs5, kSlotMs6, kSlotMs7, kSlotMs8,
  // the second kind's query inputs while the first kind's query still runs
  kSlotMsB0, kSlotMsB5, kSlotMsB6, kSlotMsB7,
  // fused step outputs (valid until the next fused call)
  kSlotOutPairs, kSlotOutRigidQ, kSlotOutRigidT, kSlotOutRigidPair, kSlotOutReflN,
  kSlotOutReflD, kSlotOutReflPair, kSlotOutMsRigidPts, kSlotOutMsReflPts, kSlotOutMsSeeds,
  kSlotOutMsRigid, kSlotOutMsRigidW, kSlotOutMsRefl, kSlotOutMsReflW,
  // batches of meshes
  kSlotBatchMeta, kSlotBatchSeg, kSlotBatchPairOff, kSlotBatchReflOff, kSlotBatchMs0,
  kSlotBatchMs1, kSlotBatchMs2, kSlotBatchMs3, kSlotBatchMs4, kSlotBatchMs5,
  // verification
  kSlotVerPoints, kSlotVerMotions, kSlotVerNbrOff,
  kSlotVerNbrIdx, kSlotVerAccepted, kSlotVerLabel, kSlotVerDist, kSlotVerPos, kSlotVerBounds,
  kSlotVerVid, kSlotVerRoot, kSlotVerD2, kSlotVerStats, kSlotVerKeys, kSlotVerKeys2,
  kSlotVerIds, kSlotVerOrder,
  kSlotPrep0, kSlotPrep1, kSlotPrep2, kSlotPrep3, kSlotPrep4, kSlotPrep5, kSlotPrep6,
  kSlotPrep7, kSlotPrep8, kSlotPrep9, kSlotPrep10, kSlotPrep11, kSlotPrep12, kSlotPrep13,
  kSlotPrep14, kSlotPrep15, kSlotPrep16, kSlotPrep17, kSlotPrep18,
  // scratch bank 1: the sort, the scan and the mean-shift set-up of a run
  // queued on the side stream beside bank-0 work
  kSlotSortB0, kSlotSortB1, kSlotSortB2, kSlotSortBHist, kSlotScanB,
  kSlotMsBankPart, kSlotMsBankKeys, kSlotMsBankVals, kSlotMsBankSeedKeys, kSlotMsBankAos,
  // pairing plan: positions at bucketed positions (FP64, FP32)
  kSlotPlanPosB, kSlotPlanPosB32, kSlotPlanPivots,
  // radix sort: digit counts, tile tickets and tile states (banks 0 and 1)
  kSlotSortState, kSlotSortBState,
  kSlotCount
};

// ------------------------------------------------------------- primitives
// Exclusive prefix sum of n uint64 values into out[0..n] (out[n] = total);
// the total also goes to *total_also when given (one launch fewer for a
// caller that reads it back).
void exclusive_scan_u64(Session *s, const uint64_t *in, uint64_t *out, uint64_t n,
                        uint64_t *total_also = nullptr);
// Exclusive prefix sum of n uint32 values into uint64 out[0..n].
void exclusive_scan_u32(Session *s, const uint32_t *in, uint64_t *out, uint64_t n,
                        uint64_t *total_also = nullptr);
// Stable LSD radix sort of (key, value) by the low `bits` bits of key.
void radix_sort_pairs_u64(Session *s, const uint64_t *keys_in, const uint32_t *vals_in,
                          uint64_t *keys_out, uint32_t *vals_out, uint64_t n, int bits);

// Zero device memory with a kernel (no memset nodes when a caller captures).
void zero_device(Session *s, void *ptr, size_t bytes);

// std::mt19937_64 step on a state in libstdc++ layout (sampling.cpp).
uint64_t mt64_next(mvs_mt64 *g);

// splitmix64 output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): seed-stratum draws.
__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__host__ __device__ inline unsigned blocks_for(uint64_t n, unsigned threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

// MI355X deals the blocks of a launch to its 8 XCDs round-robin, and each XCD
// has its own L2.  A kernel launched on xcd_grid(nb) blocks takes its logical
// block from xcd_block(nb): each XCD then works on one contiguous range of
// logical blocks, so data shared by neighbouring blocks (mesh rings) stays in
// one L2.  Logical blocks >= nb are past the end.
constexpr unsigned kXcdCount = 8;
inline unsigned xcd_grid(unsigned nb) { return (nb + kXcdCount - 1) / kXcdCount * kXcdCount; }
__device__ inline unsigned xcd_block(unsigned nb) {
  const unsigned per = (nb + kXcdCount - 1) / kXcdCount;
  return (blockIdx.x % kXcdCount) * per + blockIdx.x / kXcdCount;
}

// ------------------------------------------------------------ device math
// All FP64 helpers compile with -ffp-contract=off so every product and sum
// rounds as written, matching the reference CPU build
// (src/core/math.cpp:9-38, 113-119).
struct V3 {
  double x, y, z;
};
__host__ __device__ inline V3 v3(double x, double y, double z) { return {x, y, z}; }
__host__ __device__ inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline V3 vdivs(V3 a, double s) { return a * (1 / s); }
__host__ __device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double norm(V3 a) { return sqrt(dot(a, a)); }
__host__ __device__ inline V3 unit(V3 a) {
  double n = norm(a);
  return n > 1e-14 ? vdivs(a, n) : V3{0, 0, 0};
}
__host__ __device__ inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
struct M3 {
  double m[9];
};
__host__ __device__ inline V3 apply(const M3 &m, V3 v) {
  return {m.m[0] * v.x + m.m[1] * v.y + m.m[2] * v.z, m.m[3] * v.x + m.m[4] * v.y + m.m[5] * v.z,
          m.m[6] * v.x + m.m[7] * v.y + m.m[8] * v.z};
}
__host__ __device__ inline M3 frame(V3 a, V3 b, V3 c) {
  return {{a.x, b.x, c.x, a.y, b.y, c.y, a.z, b.z, c.z}};
}
// R = B * A^T accumulated from zero in k order (math.cpp:31-38).
__host__ __device__ inline M3 mul_bt(const M3 &b, const M3 &a) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k)
        s += b.m[i * 3 + k] * a.m[j * 3 + k];
      r.m[i * 3 + j] = s;
    }
  return r;
}
__host__ __device__ inline M3 reflection(V3 n) {
  M3 r = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.m[i * 3 + j] -= 2 * comp(n, i) * comp(n, j);
  return r;
}

// Wavefront helpers (64 lanes on CDNA).
constexpr int kWave = 64;
__device__ inline int lane_id() { return __lane_id(); }
__device__ inline uint64_t ballot(bool p) { return __ballot(p); }
__device__ inline int lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}
// Cross-lane steps through DPP: register moves between the lanes of a
// 16-lane row, with no LDS round trip (loops that run them per item).
template <int kCtrl, int kRowMask = 0xF>
__device__ inline uint32_t dpp_u32(uint32_t old, uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
      static_cast<int>(old), static_cast<int>(v), kCtrl, kRowMask, 0xF, true));
}
// Inclusive prefix sum over the 64 lanes: Hillis-Steele inside each row
// (row_shr 1, 2, 4, 8), then row 0 into row 1 and row 2 into row 3
// (row_bcast:15), then rows 0-1 into rows 2-3 (row_bcast:31).
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
  x += dpp_u32<0x111>(0u, x);
  x += dpp_u32<0x112>(0u, x);
  x += dpp_u32<0x114>(0u, x);
  x += dpp_u32<0x118>(0u, x);
  x += dpp_u32<0x142, 0xA>(0u, x);
  x += dpp_u32<0x143, 0xC>(0u, x);
  return x;
}
template <int kCtrl> __device__ inline float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kCtrl, 0xF, 0xF, true));
}
template <int kCtrl> __device__ inline double dpp_f(double v) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
  const uint32_t lo = static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(static_cast<uint32_t>(b)), kCtrl, 0xF, 0xF, true));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(
      static_cast<int>(static_cast<uint32_t>(b >> 32)), kCtrl, 0xF, 0xF, true));
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}
__device__ inline float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double lane_value(double v, int l) {
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(b), l);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(b >> 32), l);
  return __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(hi) << 32) | lo));
}
// Sum over the 64 lanes, the same (wave-uniform) value in every lane and a
// fixed order: quads (quad_perm), rows (row_ror 4, 8), then the four rows.
template <class R> __device__ inline R wave_total(R v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f<0x124>(v);  // row_ror:4
  v += dpp_f<0x128>(v);  // row_ror:8
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

template <class T> __device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, kWave);
  return v;
}

}  // namespace mvs
