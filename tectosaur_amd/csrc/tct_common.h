// Shared helpers for the libtct_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/tct_hip.h"

namespace tct {

// Record the message of a failed call for tct_last_error(); returns code.
int set_error(int code, const char *fmt, ...);

#define TCT_CHECK_ARG(cond, ...)                                 \
    do {                                                         \
        if (!(cond)) return ::tct::set_error(TCT_EARG, __VA_ARGS__); \
    } while (0)

#define TCT_CHECK_LAUNCH(what)                                                         \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return ::tct::set_error(TCT_EHIP, "%s: %s", what, hipGetErrorString(e_));  \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// 1/sqrt(x) in FP64 for finite x > 0: the hardware estimate refined by one
// third-order Newton step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, which
// takes the estimate's error below double rounding.  No special-value
// handling (3 instructions fewer than the library routine): x = 0 gives a
// non-finite value, which callers replace by 0.
__device__ __forceinline__ double rsq_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Kernel material constants (tectosaur/kernels/__init__.py constants_code).
struct ElasticConsts {
    double CsU0, CsU1;        // U
    double CsT0, CsT1;        // T, A
    double CsH0, CsH1, CsH2, CsH3, nu;  // H
    double CsRT0, CsRT1, CsRT2, CsRH0, CsRH1, CsRH2;  // regularized
};

inline ElasticConsts elastic_consts(const double *params) {
    const double G = params[0], nu = params[1];
    ElasticConsts c;
    c.CsU0 = (3.0 - 4.0 * nu) / (G * 16.0 * M_PI * (1.0 - nu));
    c.CsU1 = 1.0 / (G * 16.0 * M_PI * (1.0 - nu));
    c.CsT0 = (1 - 2.0 * nu) / (8.0 * M_PI * (1.0 - nu));
    c.CsT1 = 3.0 / (8.0 * M_PI * (1.0 - nu));
    c.CsH0 = G / (4 * M_PI * (1 - nu));
    c.CsH1 = 1 - 2 * nu;
    c.CsH2 = -1 + 4 * nu;
    c.CsH3 = 3 * nu;
    c.nu = nu;
    c.CsRT0 = -1.0 / (8.0 * M_PI * (1.0 - nu));
    c.CsRT1 = -(1 - 2 * nu) / (8.0 * M_PI * (1.0 - nu));
    c.CsRT2 = -1.0 / (4 * M_PI);
    c.CsRH0 = G / (8.0 * M_PI * (1 - nu));
    c.CsRH1 = 2 * (1 - nu);
    c.CsRH2 = 4 * nu;
    return c;
}

}  // namespace tct
