// hk_internal.h -- host-side helpers shared by the library's translation
// units: error reporting and argument checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "heliosk.h"

namespace hk {

void set_error(const char *fmt, ...);

#define HK_CHECK_ARG(cond, ...)                                   \
    do {                                                          \
        if (!(cond)) { ::hk::set_error(__VA_ARGS__); return HK_ERR_ARG; } \
    } while (0)

#define HK_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::hk::set_error("%s: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                            __FILE__, __LINE__);                              \
            return HK_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

#define HK_LAUNCH_CHECK() HK_HIP(hipGetLastError())

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// HK_VOIGT_COMPAT sum (hk_compat.hip); ws holds hk_voigt_workspace_bytes.
int compat_accumulate(const hk_line_state *st, const hk_grid *grid, double kmin, int accumulate,
                      double *K, void *ws, hipStream_t s);

}  // namespace hk
