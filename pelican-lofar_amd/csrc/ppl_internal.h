// ppl_internal.h -- shared internals of the C-ABI library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/ppl_amd.h"

namespace ppl {

// Thread-local message for ppl_last_error().
void set_error(const char* fmt, ...);

#define PPL_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::ppl::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,                \
                             hipGetErrorString(e_));                                   \
            return PPL_E_HIP;                                                          \
        }                                                                              \
    } while (0)

#define PPL_CHECK_ARG(cond, ...)                                                       \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::ppl::set_error(__VA_ARGS__);                                             \
            return PPL_E_ARG;                                                          \
        }                                                                              \
    } while (0)

// Launch-error check after a kernel launch.
#define PPL_LAUNCHED()                                                                 \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) {                                                        \
            ::ppl::set_error("%s:%d launch: %s", __FILE__, __LINE__,                   \
                             hipGetErrorString(e_));                                   \
            return PPL_E_HIP;                                                          \
        }                                                                              \
    } while (0)

inline hipStream_t stream_of(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned cdiv(unsigned a, unsigned b) { return (a + b - 1) / b; }

// Wave size on CDNA.
constexpr int kWave = 64;

// Looks up the current device's scratch of ppl_detect_noalloc (ppl_detect.hip).
int detect_noalloc_setup();

}  // namespace ppl
