// Error reporting, device queries and the ABI version of libseemps_hip.so.
#include <cstdarg>

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

}  // namespace seemps

extern "C" const char* seemps_last_error(void) { return seemps::g_last_error; }

extern "C" int seemps_abi_version(void) { return 1; }
