// Error reporting for the C ABI.
#include <cstdarg>

#include "common.h"

namespace mt {

static thread_local char g_last_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace mt

extern "C" const char* mt_get_last_error_string(void) { return mt::g_last_error; }

extern "C" int32_t mt_version(void) { return 100; }
