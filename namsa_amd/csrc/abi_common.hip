// Error reporting and device queries shared by every entry point.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "namsa_hip.h"

namespace namsa {

static thread_local char g_last_error[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t err, const char* what) {
  if (err == hipSuccess) return kOk;
  return fail(kErrHip, "%s: %s", what, hipGetErrorString(err));
}

int device_cu_count() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (dev >= static_cast<int>(cache.size())) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}

}  // namespace namsa

extern "C" {

int namsa_abi_version(void) { return NAMSA_ABI_VERSION; }

const char* namsa_last_error(void) { return namsa::g_last_error; }

int namsa_init(int device) {
  int n = 0;
  int rc = namsa::check_hip(hipGetDeviceCount(&n), "hipGetDeviceCount");
  if (rc) return rc;
  if (device < 0 || device >= n)
    return namsa::fail(namsa::kErrArg, "namsa_init: device %d out of range (%d devices)", device, n);
  rc = namsa::check_hip(hipSetDevice(device), "hipSetDevice");
  if (rc) return rc;
  (void)namsa::device_cu_count();
  return namsa::scratch_setup();
}

}  // extern "C"
