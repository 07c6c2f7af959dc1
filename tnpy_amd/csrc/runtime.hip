// Library-wide state: the thread-local error message and the version.
#include "tnx_common.h"

#include <cstdarg>

namespace tnx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

}  // namespace tnx

extern "C" {

const char* tnx_last_error(void) { return tnx::last_error(); }

int tnx_version(void) { return 100; }

}  // extern "C"
