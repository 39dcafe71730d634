// Last-error text of the C-ABI (per thread).
#include <cstdio>
#include <cstring>

#include "host_common.hpp"

static thread_local char g_last_error[512] = "";

int plbfs_fail(int rc, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return rc;
}

extern "C" const char *plbfs_last_error(void) { return g_last_error; }
