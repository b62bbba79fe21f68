// Library-wide entries: version and per-thread error text.
#include <cstdarg>

#include "tnh_common.h"

namespace tnh {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace tnh

extern "C" {

int tnh_version(void) { return 100; }

const char* tnh_last_error(void) { return tnh::g_err; }

}  // extern "C"
