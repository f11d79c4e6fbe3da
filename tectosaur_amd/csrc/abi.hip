// Error reporting and version of the C ABI.
#include "tct_common.h"

#include <cstring>

namespace tct {

static thread_local char g_last_error[512] = "";

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace tct

extern "C" const char *tct_last_error(void) { return tct::g_last_error; }

extern "C" int tct_abi_version(void) { return 1; }
