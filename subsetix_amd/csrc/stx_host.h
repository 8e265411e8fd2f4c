// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Library-private: shared between the translation units of the library but
// not exported (the exported surface is exactly include/subsetix_hip.h).
#define STX_INTERNAL __attribute__((visibility("hidden")))

extern "C" {
// Records a message for stx_last_error() and returns `code`.
STX_INTERNAL int32_t stx_fail(int32_t code, const char* msg);
// Checks hipGetLastError() after a launch; returns 0 or STX_EHIP.
STX_INTERNAL int32_t stx_launch_status(const char* what);
// Bound on look-back polls before a block gives up (see stx_common.h).
STX_INTERNAL int64_t stx_spin_limit(void);
}
