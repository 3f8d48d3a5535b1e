// Error reporting and version of the C ABI.
#include "common.h"
#include <string.h>

namespace dlp {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace dlp

extern "C" const char* dlp_last_error(void) { return dlp::g_err; }

extern "C" int dlp_abi_version(void) { return 2; }
