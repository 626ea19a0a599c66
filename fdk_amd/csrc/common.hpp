// Shared host-side plumbing for the fdk_amd library: error state, status
// macros and small device-memory RAII.  No reference counterpart beyond the
// reference's VERIFY_TRUE / CUDA_CHECK_ERROR (common/type.h:194-196,
// common/cudaLib.cuh:49-56), which throw; this library returns status codes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "fdk_amd.h"

namespace fdk {

// Word 3 of a buffer resource for raw (untyped) dword access on gfx9 / CDNA:
// data format 32, no swizzle.  Loads past the resource's byte size return 0.
constexpr int kRawRsrcWord3 = 0x00020000;

// Records a printf-style message as the calling thread's last error and
// returns code, so call sites can `return fail(...)`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Stream-ordered allocation from the library's per-device staging pool
// (release with hipFreeAsync on a stream).  Freed blocks stay mapped in the
// pool until the process ends, so a batch-sized staging buffer is mapped once,
// not on every call.
int stage_alloc(void** p, size_t bytes, hipStream_t st);

}  // namespace fdk

#define FDK_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t fdk_e_ = (call);                                                        \
    if (fdk_e_ != hipSuccess)                                                          \
      return ::fdk::fail(fdk_e_ == hipErrorOutOfMemory ? FDK_ERR_OOM : FDK_ERR_HIP,    \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(fdk_e_),    \
                         __FILE__, __LINE__);                                          \
  } while (0)

#define FDK_REQUIRE(cond, ...)                                \
  do {                                                        \
    if (!(cond)) return ::fdk::fail(FDK_ERR_INVALID, __VA_ARGS__); \
  } while (0)
