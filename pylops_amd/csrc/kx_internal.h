// kx_internal.h -- host-side helpers shared by the translation units of
// libkirchx.so (hidden symbols, not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "kirchx.h"

// Records a printf-style message for kx_last_error() and returns code.
extern "C" __attribute__((visibility("hidden"))) int kx_failf(int code, const char *fmt, ...);
// Validates a plan (KX_OK or an error code with its message recorded).
extern "C" __attribute__((visibility("hidden"))) int kx_plan_check(const kx_plan *p);

// Returned (with no message) when a plan's forward cannot convolve its
// traces inside the spread launch; the caller then spreads and convolves
// in two launches.
#define KX_ENOFUSE (-100)

// The forward operator in one launch: the spread of m convolved with the
// plan's wavelet into d, or KX_ENOFUSE.
extern "C" __attribute__((visibility("hidden"))) int kx_spread_wavelet(const kx_plan *p,
                                                                      const void *m, void *d,
                                                                      void *stream);

#define KX_LAUNCHED(name)                                                             \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess)                                                         \
            return kx_failf((int)e_, "%s launch: %s", name, hipGetErrorString(e_));  \
    } while (0)
