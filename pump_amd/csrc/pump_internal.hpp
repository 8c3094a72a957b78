// pump_internal.hpp -- shared host-side helpers of the HIP entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pump_abi.h"

namespace pump {

inline hipStream_t as_stream(pump_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Records the failing HIP call for pump_last_error() (thread-local).
void note_hip_error(hipError_t e, const char *where, int line);

// Returns PUMP_E_HIP from the enclosing entry point on a HIP error.
#define PUMP_HIP(expr)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            ::pump::note_hip_error(e_, __FILE__, __LINE__);                                        \
            return PUMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

// Launch check: catches configuration errors of the preceding launch.
#define PUMP_LAUNCHED()                                                                            \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) {                                                                    \
            ::pump::note_hip_error(e_, __FILE__, __LINE__);                                        \
            return PUMP_E_HIP;                                                                     \
        }                                                                                          \
    } while (0)

#define PUMP_REQUIRE(cond)                                                                         \
    do {                                                                                           \
        if (!(cond))                                                                               \
            return PUMP_E_INVALID;                                                                 \
    } while (0)

// Grid size for a flat launch over n items, grid-stride beyond the cap.
inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    if (g > cap)
        g = cap;
    return (unsigned)g;
}

} // namespace pump
