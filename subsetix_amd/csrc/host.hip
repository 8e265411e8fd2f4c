// C-ABI housekeeping: error strings, launch checks, tunables.
#include <stdio.h>
#include <string.h>

#include "stx_common.h"
#include "stx_host.h"

static thread_local char g_last_error[512] = "";
static int64_t g_spin_limit = 1ll << 24;

extern "C" {

int32_t stx_fail(int32_t code, const char* msg) {
    snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
    return code;
}

int32_t stx_launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
        return stx::STX_EHIP;
    }
    return stx::STX_OK;
}

int64_t stx_spin_limit(void) { return g_spin_limit; }

const char* stx_last_error(void) { return g_last_error; }

void stx_set_spin_limit(int64_t polls) { g_spin_limit = polls > 0 ? polls : 1; }

int32_t stx_abi_version(void) { return 1; }

}  // extern "C"
