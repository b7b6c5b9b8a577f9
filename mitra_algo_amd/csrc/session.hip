// Session lifetime, error transport, scratch arenas, stream and timing.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"
#include "session.hpp"

namespace mvs {

namespace {
thread_local std::string g_error;
thread_local int g_code = MVS_OK;
// per-thread device-memory and host-copy counters (mvs_thread_metrics)
thread_local mvs_memory_metrics g_mem{};
}  // namespace

void note_device_alloc(size_t bytes) {
  g_mem.live_bytes += bytes;
  if (g_mem.live_bytes > g_mem.peak_bytes)
    g_mem.peak_bytes = g_mem.live_bytes;
}
void note_device_free(size_t bytes) { g_mem.live_bytes -= bytes; }
void note_copy(bool to_device, size_t bytes) {
  (to_device ? g_mem.h2d_bytes : g_mem.d2h_bytes) += bytes;
}

void set_error(int code, const char *msg) {
  g_code = code;
  g_error = msg;
}

Session *as_session(mvs_session *s) {
  if (!s)
    raise(MVS_EINVAL, "null session");
  return reinterpret_cast<Session *>(s);
}

hipStream_t session_stream(Session *s) { return s->stream; }

void *scratch(Session *s, int slot, size_t bytes) {
  if (Arena *a = s->arena) {
    const size_t want = bytes ? bytes : 1;
    if (a->len[slot] >= want)
      return a->base + a->off[slot];
    const size_t at = (a->used + 255) / 256 * 256;
    if (at > a->cap || want > a->cap - at)
      raise(MVS_EINVAL, "the allocation-free call needs more scratch than its static buffer");
    a->off[slot] = at;
    a->len[slot] = want;
    a->used = at + want;
    return a->base + at;
  }
  auto &b = s->slots[slot];
  if ((bytes ? bytes : 1) > b.n) {
    if (s->replay == Replay::kCapture)
      raise(MVS_EINVAL, "graph capture needs the buffers of its recorded step");
    s->drop_graph();  // a captured graph points at the old allocation
    // queued work may still read the old allocation
    MVS_HIP(hipStreamSynchronize(s->stream));
    b.alloc(bytes ? bytes : 1);
  }
  return b.ptr;
}

void *host_stage(Session *s, size_t bytes) {
  if (bytes > s->stage_bytes) {
    // queued copies may still use the old staging buffer
    MVS_HIP(hipStreamSynchronize(s->stream));
    if (s->stage)
      (void)hipHostFree(s->stage);
    s->stage = nullptr;
    s->stage_bytes = 0;
    MVS_HIP(hipHostMalloc(&s->stage, bytes, hipHostMallocDefault));
    s->stage_bytes = bytes;
  }
  return s->stage;
}

// Host <-> device copies are counted, and timed under "copy_h2d" /
// "copy_d2h" when the session times its launches.
void copy_h2d(Session *s, void *dst, const void *src, size_t bytes) {
  if (!bytes)
    return;
  s->begin("copy_h2d");
  MVS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream));
  s->end();
  note_copy(true, bytes);
}
void copy_d2h(Session *s, void *dst, const void *src, size_t bytes) {
  if (!bytes)
    return;
  s->begin("copy_d2h");
  MVS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream));
  s->end();
  note_copy(false, bytes);
  MVS_HIP(hipStreamSynchronize(s->stream));
}
void sync(Session *s) { MVS_HIP(hipStreamSynchronize(s->stream)); }

void Session::begin(const char *name) {
  if (timing_open)  // the previous launch raised before closing its record
    end();
  if (!timing)
    return;
  if (!timing_only.empty() && timing_only.find("," + std::string(name) + ",") == std::string::npos)
    return;
  hipEvent_t a, b;
  if (event_pool.size() >= 2) {
    a = event_pool.back();
    event_pool.pop_back();
    b = event_pool.back();
    event_pool.pop_back();
  } else {
    MVS_HIP(hipEventCreate(&a));
    MVS_HIP(hipEventCreate(&b));
  }
  MVS_HIP(hipEventRecord(a, stream));
  pending.push_back({name, a, b});
  timing_open = true;
}
void Session::end() {
  if (!timing_open)
    return;
  timing_open = false;
  MVS_HIP(hipEventRecord(pending.back().stop, stream));
}
void Session::collect() {
  if (pending.empty())
    return;
  MVS_HIP(hipStreamSynchronize(stream));
  for (auto &p : pending) {
    float ms = 0;
    MVS_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    auto &acc = timings[p.name];
    acc.first += ms;
    acc.second += 1;
    event_pool.push_back(p.start);
    event_pool.push_back(p.stop);
  }
  pending.clear();
}
void Session::check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    raise(MVS_EHIP, std::string("HIP: launch ") + what + ": " + hipGetErrorString(e));
}

void Session::drop_graph() {
  if (graph_exec) {
    (void)hipStreamSynchronize(own_stream);
    (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
  }
}

Session::~Session() {
  drop_graph();
  for (auto e : graph_events)
    if (e)
      (void)hipEventDestroy(e);
  for (auto &p : pending) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : event_pool)
    (void)hipEventDestroy(e);
  if (stage)
    (void)hipHostFree(stage);
  if (batch_stage)
    (void)hipHostFree(batch_stage);
  if (side_stream) {
    (void)hipStreamSynchronize(side_stream);
    (void)hipStreamDestroy(side_stream);
  }
  for (auto e : {fork_event, join_event, setup_events[0], setup_events[1]})
    if (e)
      (void)hipEventDestroy(e);
  if (own_stream)
    (void)hipStreamDestroy(own_stream);
}

}  // namespace mvs

using namespace mvs;

extern "C" {

const char *mvs_last_error(void) { return g_error.c_str(); }

int mvs_session_create(int device, mvs_session **out) {
  return guarded([&] {
    if (!out)
      raise(MVS_EINVAL, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      raise(MVS_ENODEV, "HIP: no available device");
    if (device < 0 || device >= count)
      raise(MVS_EINVAL, "device ordinal out of range");
    MVS_HIP(hipSetDevice(device));
    auto *s = new Session();
    s->device = device;
    hipError_t e = hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete s;
      hip_check(e, "hipStreamCreateWithFlags");
    }
    s->stream = s->own_stream;
    *out = reinterpret_cast<mvs_session *>(s);
  });
}

void mvs_session_destroy(mvs_session *s) {
  if (!s)
    return;
  auto *p = reinterpret_cast<Session *>(s);
  (void)hipSetDevice(p->device);
  (void)hipStreamSynchronize(p->stream);
  delete p;
}

int mvs_session_set_stream(mvs_session *s, void *stream) {
  return guarded([&] {
    Session *p = as_session(s);
    p->stream = stream == MVS_OWN_STREAM ? p->own_stream : static_cast<hipStream_t>(stream);
  });
}

int mvs_session_sync(mvs_session *s) {
  return guarded([&] { sync(as_session(s)); });
}

int mvs_dev_copy(mvs_session *sp, void *dst, const void *src, uint64_t bytes) {
  return guarded([&] {
    Session *s = as_session(sp);
    if (bytes)
      MVS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->stream));
  });
}

int mvs_session_enable_timing(mvs_session *s, int enable) {
  return guarded([&] {
    Session *p = as_session(s);
    p->collect();
    p->timing = enable != 0;
    p->timings.clear();
  });
}

int mvs_session_timing_filter(mvs_session *s, const char *names) {
  return guarded([&] {
    Session *p = as_session(s);
    p->collect();
    p->timing_only = names && *names ? "," + std::string(names) + "," : std::string();
  });
}

int mvs_dev_timing(mvs_session *s, const char *kernel, double *ms, uint64_t *launches) {
  return guarded([&] {
    Session *p = as_session(s);
    p->collect();
    double t = 0;
    uint64_t n = 0;
    for (const auto &kv : p->timings)
      if (kernel ? kv.first == kernel : kv.first.rfind("copy_", 0) != 0) {
        t += kv.second.first;
        n += kv.second.second;
      }
    if (ms)
      *ms = t;
    if (launches)
      *launches = n;
  });
}

int mvs_thread_metrics(mvs_memory_metrics *out) {
  if (!out) {
    set_error(MVS_EINVAL, "null argument");
    return MVS_EINVAL;
  }
  *out = g_mem;
  return MVS_OK;
}

void mvs_thread_metrics_reset(void) {
  g_mem.peak_bytes = g_mem.live_bytes;
  g_mem.h2d_bytes = g_mem.d2h_bytes = 0;
}

}  // extern "C"
