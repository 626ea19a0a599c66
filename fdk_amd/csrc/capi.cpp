// Error state and device queries of the C ABI (include/fdk_amd.h).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
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

int stage_alloc(void** p, size_t bytes, hipStream_t st) {
  static std::mutex mu;
  static std::map<int, hipMemPool_t> pools;
  int dev = 0;
  FDK_HIP(hipGetDevice(&dev));
  hipMemPool_t pool = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = pools.find(dev);
    if (it != pools.end()) {
      pool = it->second;
    } else {
      hipMemPoolProps props{};
      props.allocType = hipMemAllocationTypePinned;
      props.handleTypes = hipMemHandleTypeNone;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      FDK_HIP(hipMemPoolCreate(&pool, &props));
      uint64_t keep = UINT64_MAX;  // freed blocks stay mapped for the next call
      hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) {
        (void)hipMemPoolDestroy(pool);
        return fail(FDK_ERR_HIP, "staging pool: %s", hipGetErrorString(e));
      }
      pools[dev] = pool;
    }
  }
  FDK_HIP(hipMallocFromPoolAsync(p, bytes, pool, st));
  return FDK_OK;
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
