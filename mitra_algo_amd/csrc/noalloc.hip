// Allocation-free batched mean-shift: mvs_dev_ms_batched whose scratch
// (grid control block and moments, cell keys, the radix sort's buffers, the
// cell-sorted records, cell offsets, the seeds' order) comes from a static
// device array instead of the session's grow-only slots, so a call makes no
// hipMalloc / hipFree and no pinned host allocation.  For callers that cannot
// allocate (code under graph capture, tight serving loops).  No counterpart in
// the reference.
//
// Each GPU holds its own copy of `scratch` for the life of the process; its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and kept.  Every call on a GPU uses that GPU's array, which
// keeps the previous call's contents: what must start at zero (the control
// block) is cleared by a fill kernel.  Calls that overlap on one GPU (two
// sessions, two threads) race on the array.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "pairing.hpp"

namespace {
// 16 MiB: the scratch of MVS_NOALLOC_MAX_POINTS points and
// MVS_NOALLOC_MAX_SEEDS seeds in FP64 (about 11.5 MiB), with room to spare.
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace mvs {
namespace {

std::mutex g_lock;
void *g_scratch[kMaxDevices] = {};

// The device address of this device's `scratch` (the host-side name is only a
// stand-in for it), looked up on the first call with the device current.
void *device_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    raise(MVS_ENODEV, "HIP: device index out of range");
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_scratch[device]) {
    int cur = 0;
    MVS_HIP(hipGetDevice(&cur));
    if (cur != device)
      MVS_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (cur != device)
      (void)hipSetDevice(cur);
    hip_check(e, "hipGetSymbolAddress");
    g_scratch[device] = ptr;
  }
  return g_scratch[device];
}

template <class R>
void ms_batched_noalloc(Session *s, int dim, uint64_t n, const R *pts, uint64_t ns,
                        const R *seeds, int iterations, double h, double cutoff, R *out,
                        R *weight) {
  if (n > MVS_NOALLOC_MAX_POINTS || ns > MVS_NOALLOC_MAX_SEEDS)
    raise(MVS_EINVAL, "the allocation-free mean-shift takes at most "
                      "MVS_NOALLOC_MAX_POINTS points and MVS_NOALLOC_MAX_SEEDS seeds");
  Arena arena;
  arena.base = static_cast<unsigned char *>(device_scratch(s->device));
  arena.cap = kScratchBytes;
  const bool timing = s->timing;
  struct Restore {
    Session *s;
    bool timing;
    ~Restore() {
      s->arena = nullptr;
      s->timing = timing;
    }
  } restore{s, timing};
  s->arena = &arena;
  s->timing = false;  // no events either
  ms_begin(s);
  if constexpr (sizeof(R) == 4)
    ms_batched_f32(s, dim, n, pts, ns, seeds, nullptr, 0, iterations, h, cutoff, out, weight,
                   -1);
  else
    ms_batched_f64(s, dim, n, pts, ns, seeds, nullptr, 0, iterations, h, cutoff, out, weight,
                   -1);
  // the non-finite-input flag: a synchronous 4-byte read into a local
  unsigned bad = 0;
  MVS_HIP(hipMemcpyAsync(&bad, ms_bad_flag(s), sizeof bad, hipMemcpyDeviceToHost, s->stream));
  MVS_HIP(hipStreamSynchronize(s->stream));
  if (bad)
    raise(MVS_EINVAL, "non-finite mean-shift points");
}

}  // namespace
}  // namespace mvs

using namespace mvs;

extern "C" {

int mvs_dev_ms_batched_noalloc(mvs_session *sp, int precision, int dim, uint64_t n,
                               const void *pts, uint64_t ns, const void *seeds, int iterations,
                               double h, double cutoff, void *out, void *weight) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (precision == MVS_FP32_GUARDED)
      ms_batched_noalloc<float>(s, dim, n, static_cast<const float *>(pts), ns,
                                static_cast<const float *>(seeds), iterations, h, cutoff,
                                static_cast<float *>(out), static_cast<float *>(weight));
    else if (precision == MVS_FP64)
      ms_batched_noalloc<double>(s, dim, n, static_cast<const double *>(pts), ns,
                                 static_cast<const double *>(seeds), iterations, h, cutoff,
                                 static_cast<double *>(out), static_cast<double *>(weight));
    else
      raise(MVS_EINVAL, "unknown precision");
  });
}

}  // extern "C"
