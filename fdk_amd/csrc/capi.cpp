// Error state and device queries of the C ABI (include/fdk_amd.h).
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace fdk {

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

}  // namespace fdk

extern "C" int fdk_last_error(char* buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return (int)g_last_error.size();
}

extern "C" int fdk_version(void) { return 100; }

extern "C" int fdk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// Device buffers for the C++ layer's DevData / DevArray3D (cudaBP/cudaLib.cuh:298-420,
// 552-640): linear allocations on the current device, copies in either direction.
extern "C" int fdk_dev_malloc(void** out, size_t bytes) {
  FDK_REQUIRE(out, "null output pointer");
  *out = nullptr;
  if (bytes == 0) return FDK_OK;
  FDK_HIP(hipMalloc(out, bytes));
  return FDK_OK;
}

extern "C" int fdk_dev_free(void* p) {
  if (p) FDK_HIP(hipFree(p));
  return FDK_OK;
}

extern "C" int fdk_memcpy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return FDK_OK;
  FDK_REQUIRE(dst && src, "null pointer");
  FDK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  return FDK_OK;
}

extern "C" int fdk_memcpy_2d(void* dst, size_t dst_pitch, const void* src, size_t src_pitch,
                             size_t width_bytes, size_t rows) {
  if (width_bytes == 0 || rows == 0) return FDK_OK;
  FDK_REQUIRE(dst && src, "null pointer");
  FDK_REQUIRE(dst_pitch >= width_bytes && src_pitch >= width_bytes, "pitch below row width");
  FDK_HIP(hipMemcpy2D(dst, dst_pitch, src, src_pitch, width_bytes, rows, hipMemcpyDefault));
  return FDK_OK;
}

extern "C" int fdk_dev_memset(void* p, int value, size_t bytes) {
  if (bytes == 0) return FDK_OK;
  FDK_REQUIRE(p, "null pointer");
  FDK_HIP(hipMemset(p, value, bytes));
  return FDK_OK;
}
