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
