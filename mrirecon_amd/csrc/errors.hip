// Thread-local error message of the C ABI.
#include "mrh_common.h"
#include "mrirecon_hip.h"

static thread_local char g_err[1024] = "";

void mrh_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* mrh_last_error(void) { return g_err; }

extern "C" int mrh_abi_version(void) { return 13; }
