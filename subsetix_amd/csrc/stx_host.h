// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" {
// Records a message for stx_last_error() and returns `code`.
int32_t stx_fail(int32_t code, const char* msg);
// Checks hipGetLastError() after a launch; returns 0 or STX_EHIP.
int32_t stx_launch_status(const char* what);
// Bound on look-back polls before a block gives up (see stx_common.h).
int64_t stx_spin_limit(void);
}
