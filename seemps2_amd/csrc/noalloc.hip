// seemps_zdotc without a workspace argument, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops): the
// partial sums of the two-pass reduction go to a file-scope device array.
#include <mutex>

#include "common.h"

namespace seemps {
namespace {

// Largest number of partials reduce_blocks() produces (it caps there).
constexpr int64_t NOALLOC_PARTIALS = 512;
static __device__ __align__(256) double2 scratch[NOALLOC_PARTIALS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current (the array's name is only a host-side
// stand-in; every use in host code goes through this pointer).
int scratch_for_current_device(double2** out) {
  int dev = 0;
  SEEMPS_HIP(hipGetDevice(&dev));
  SEEMPS_CHECK_ARG(dev >= 0 && dev < MAX_DEVICES, "zdotc_noalloc: device index %d too large",
                   dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SEEMPS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double2*)p;
  }
  *out = g_scratch_ptr[dev];
  return SEEMPS_OK;
}

}  // namespace
}  // namespace seemps

extern "C" int seemps_zdotc_noalloc(int64_t n, const void* x, const void* y, void* result,
                                    void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(n >= 0, "zdotc_noalloc: negative length");
  if (n == 0) {
    return zaxpby_impl(1, make_double2(0.0, 0.0), nullptr, 0, make_double2(0.0, 0.0),
                       (double2*)result, (hipStream_t)stream);
  }
  const int64_t parts = reduce_blocks(n);
  SEEMPS_CHECK_ARG(parts <= NOALLOC_PARTIALS,
                   "zdotc_noalloc: %lld partial sums exceed the %lld of the static scratch",
                   (long long)parts, (long long)NOALLOC_PARTIALS);
  double2* work = nullptr;
  const int st = scratch_for_current_device(&work);
  if (st) return st;
  return zmdotc_impl(n, 1, (const double2*)x, n, (const double2*)y, (double2*)result, work,
                     (hipStream_t)stream);
}

