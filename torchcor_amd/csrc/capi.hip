// Error text and version of the C ABI.
#include "tc_internal.h"

namespace tc {
namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

}  // namespace tc

extern "C" const char* tc_last_error(void) { return tc::g_last_error.c_str(); }

extern "C" int tc_abi_version(void) { return 1; }
