// Library-wide C ABI helpers: thread-local error message and ABI version.
#include "common.h"
#include "cvqpg.h"

namespace cvqpg {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_err; }

}  // namespace cvqpg

extern "C" const char* cvqpg_last_error(void) { return cvqpg::get_error(); }

extern "C" int cvqpg_abi_version(void) { return 2; }
