// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <limits>

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

// Smallest s >= 0 whose correctly rounded square root is >= t (host side,
// binary search over the non-negative bit patterns); NaN for a NaN t.
template <typename T, typename Bits>
inline T sqrt_threshold(T t) {
    if (t != t) return t;
    if (!(t > (T)0)) return (T)0;
    Bits lo = 0, hi;
    const T inf = std::numeric_limits<T>::infinity();
    __builtin_memcpy(&hi, &inf, sizeof(T));
    while (lo < hi) {
        const Bits mid = lo + (hi - lo) / 2;
        T v;
        __builtin_memcpy(&v, &mid, sizeof(T));
        if (std::sqrt(v) >= t) hi = mid;
        else lo = mid + 1;
    }
    T r;
    __builtin_memcpy(&r, &lo, sizeof(T));
    return r;
}

