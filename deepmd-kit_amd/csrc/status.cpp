// Thread-local error state behind dpk_last_error().
#include <cstdarg>
#include <cstdio>

#include "dpk_common.h"

namespace dpk {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

}  // namespace dpk

extern "C" const char* dpk_last_error(void) { return dpk::g_err; }

extern "C" int dpk_abi_version(void) { return 2; }
