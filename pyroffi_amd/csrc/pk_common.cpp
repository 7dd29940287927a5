// Error reporting and version query of the C ABI.
#include "pk_common.h"

#include <cstdarg>
#include <cstdio>

namespace {
thread_local char g_last_error[512] = "";
}

namespace pk {

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace pk

extern "C" const char* pk_last_error(void) { return g_last_error; }

extern "C" int32_t pk_abi_version(void) { return 1; }
