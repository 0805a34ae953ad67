// Error reporting, device queries and the ABI version of libseemps_hip.so.
#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace seemps {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("launch of %s failed: %s", what, hipGetErrorString(e));
    return SEEMPS_ERR_RUNTIME;
  }
  return SEEMPS_OK;
}

// Pinned host staging for the small device-to-host reads of the hot paths
// (eigensolver exits, split ranks, SVD control blocks): copies into pinned
// memory are plain DMA, where copies into pageable memory go through a
// runtime bounce buffer.  One buffer per host thread, allocated on first use;
// a request larger than the buffer replaces it.  Every buffer is listed in a
// registry so that seemps_release() can free them all; it bumps a generation
// number, and a thread whose buffer belongs to an older generation allocates
// a new one on its next use.
namespace {
std::mutex g_pin_mu;
std::vector<void*> g_pin_all;
std::atomic<int> g_pin_gen{0};
}  // namespace

void* pinned_stage(size_t bytes) {
  static thread_local void* buf = nullptr;
  static thread_local size_t cap = 0;
  static thread_local int gen = -1;
  const int now = g_pin_gen.load(std::memory_order_acquire);
  if (gen != now) {  // released (or never allocated) since this thread's last use
    buf = nullptr;
    cap = 0;
    gen = now;
  }
  if (bytes > cap) {
    const size_t want = std::max<size_t>(bytes, 4096);
    void* p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_pin_mu);
    if (buf) {
      (void)hipHostFree(buf);
      g_pin_all.erase(std::remove(g_pin_all.begin(), g_pin_all.end(), buf), g_pin_all.end());
    }
    g_pin_all.push_back(p);
    buf = p;
    cap = want;
  }
  return buf;
}

int device_info(DeviceInfo* info) {
  // One cache entry per device; the library is used from one process per GPU.
  static thread_local int cached_dev = -1;
  static thread_local int cached_cu = 0;
  int dev = 0;
  SEEMPS_HIP(hipGetDevice(&dev));
  if (dev != cached_dev) {
    int cu = 0;
    SEEMPS_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    cached_dev = dev;
    cached_cu = cu;
  }
  info->device = cached_dev;
  info->num_cu = cached_cu;
  return SEEMPS_OK;
}

// ---------------------------------------------------------------------------
// Kernel timing registry.
namespace {

struct TimingSlot {
  std::string name;
  int64_t launches = 0;
  double ms = 0.0;
  double flops = 0.0;
};
struct PendingTiming {
  int slot;
  hipEvent_t start, stop;
  double flops;
};

std::atomic<bool> g_timing_on{false};
std::mutex g_timing_mu;
bool g_timing_all = false;
std::vector<std::string> g_timing_filter;
std::vector<TimingSlot> g_slots;
std::vector<PendingTiming> g_pending;
std::vector<hipEvent_t> g_free_events;  // drained events, reused by later launches

// Caller holds g_timing_mu.  nullptr if no event can be created.
hipEvent_t take_event() {
  if (!g_free_events.empty()) {
    hipEvent_t e = g_free_events.back();
    g_free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// Caller holds g_timing_mu.
int timing_slot(const char* name) {
  if (!g_timing_all) {
    bool wanted = false;
    for (const auto& f : g_timing_filter) wanted = wanted || (f == name);
    if (!wanted) return -1;
  }
  for (size_t i = 0; i < g_slots.size(); ++i)
    if (g_slots[i].name == name) return (int)i;
  TimingSlot s;
  s.name = name;
  g_slots.push_back(s);
  return (int)g_slots.size() - 1;
}

// Caller holds g_timing_mu.
int drain_pending() {
  int st = SEEMPS_OK;
  for (auto& p : g_pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.stop) == hipSuccess &&
        hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      g_slots[p.slot].launches += 1;
      g_slots[p.slot].ms += ms;
      g_slots[p.slot].flops += p.flops;
    } else {
      set_error("kernel timing: event query failed");
      st = SEEMPS_ERR_RUNTIME;
    }
    g_free_events.push_back(p.start);
    g_free_events.push_back(p.stop);
  }
  g_pending.clear();
  return st;
}

}  // namespace

TimedLaunch::TimedLaunch(const char* name, hipStream_t stream, double flops)
    : slot_(-1), start_(nullptr), stream_(stream), flops_(flops) {
  if (!g_timing_on.load(std::memory_order_relaxed)) return;
  // no event nodes inside a HIP graph being captured
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream_, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
    return;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  const int slot = timing_slot(name);
  if (slot < 0) return;
  hipEvent_t e = take_event();
  if (!e) return;
  if (hipEventRecord(e, stream_) != hipSuccess) {
    g_free_events.push_back(e);
    return;
  }
  start_ = e;
  slot_ = slot;
}

TimedLaunch::~TimedLaunch() {
  if (slot_ < 0) return;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  hipEvent_t stop = take_event();
  if (!stop) {
    g_free_events.push_back(start_);
    return;
  }
  if (hipEventRecord(stop, stream_) != hipSuccess) {
    g_free_events.push_back(start_);
    g_free_events.push_back(stop);
    return;
  }
  g_pending.push_back(PendingTiming{slot_, start_, stop, flops_});
  // bound the number of live events on long runs
  if (g_pending.size() >= 8192) (void)drain_pending();
}

}  // namespace seemps

namespace seemps {
namespace {
__global__ void timing_empty_kernel() {}
}  // namespace
}  // namespace seemps

extern "C" const char* seemps_last_error(void) { return seemps::g_last_error; }

extern "C" int seemps_timing_calibrate(int launches, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(launches >= 0, "timing_calibrate: negative count");
  for (int i = 0; i < launches; ++i) {
    TimedLaunch timed("timing_calibration", (hipStream_t)stream, 0.0);
    hipLaunchKernelGGL(timing_empty_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream);
    const int st = check_launch("timing_calibration");
    if (st) return st;
  }
  return SEEMPS_OK;
}

extern "C" int seemps_release(void) {
  using namespace seemps;
  int st = SEEMPS_OK;
  {
    std::lock_guard<std::mutex> lock(g_timing_mu);
    st = drain_pending();
    for (hipEvent_t e : g_free_events) (void)hipEventDestroy(e);
    g_free_events.clear();
    g_slots.clear();
    g_timing_on.store(false);
  }
  {
    std::lock_guard<std::mutex> lock(g_pin_mu);
    for (void* p : g_pin_all) (void)hipHostFree(p);
    g_pin_all.clear();
    g_pin_gen.fetch_add(1, std::memory_order_acq_rel);
  }
  return st;
}

extern "C" int seemps_abi_version(void) { return 1; }

extern "C" int seemps_timing_enable(const char* filter) {
  using namespace seemps;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  const int st = drain_pending();
  g_slots.clear();
  g_timing_filter.clear();
  g_timing_all = false;
  if (!filter || !*filter) {
    g_timing_on.store(false);
    return st;
  }
  std::string f(filter);
  size_t pos = 0;
  while (pos <= f.size()) {
    size_t comma = f.find(',', pos);
    if (comma == std::string::npos) comma = f.size();
    std::string item = f.substr(pos, comma - pos);
    if (item == "*") g_timing_all = true;
    if (!item.empty()) g_timing_filter.push_back(item);
    pos = comma + 1;
  }
  g_timing_on.store(true);
  return st;
}

extern "C" int seemps_timing_count(void) {
  using namespace seemps;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  (void)drain_pending();
  return (int)g_slots.size();
}

extern "C" int seemps_timing_get(int index, char* name, int64_t name_len, int64_t* launches,
                                 double* total_ms, double* flops) {
  using namespace seemps;
  std::lock_guard<std::mutex> lock(g_timing_mu);
  int st = drain_pending();
  if (st) return st;
  SEEMPS_CHECK_ARG(index >= 0 && index < (int)g_slots.size(), "timing: index out of range");
  const TimingSlot& s = g_slots[index];
  if (name && name_len > 0) {
    const size_t n = std::min((size_t)(name_len - 1), s.name.size());
    memcpy(name, s.name.data(), n);
    name[n] = '\0';
  }
  if (launches) *launches = s.launches;
  if (total_ms) *total_ms = s.ms;
  if (flops) *flops = s.flops;
  return SEEMPS_OK;
}
