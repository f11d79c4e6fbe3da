// VALU issue-rate calibration kernels (measurement only; no reference
// counterpart).  Each lane runs 16 independent pairs of dependent FMA
// chains, so issue, not latency, bounds the rate; the host times a launch
// and divides the FMA count by it.  Two operand forms per precision:
//   VGPR: every operand in a VGPR -- gfx950 issues two such instructions of
//         one SIMD in the same cycle;
//   SGPR: one operand is a wave-uniform kernel argument held in an SGPR --
//         such instructions are issued one per cycle.
// The FP32 rate with VGPR operands is the FP32 vector peak of the chip
// without packed instructions.
#include "tct_common.h"

namespace tct {

constexpr int CAL_CHAINS = 16;

template <typename T, bool SGPR>
__global__ __launch_bounds__(256) void calib_fma_kernel(T s, int iters, float *__restrict__ out) {
    T a[CAL_CHAINS], b[CAL_CHAINS], y[CAL_CHAINS];
    const T t = (T)threadIdx.x;
#pragma unroll
    for (int k = 0; k < CAL_CHAINS; ++k) {
        a[k] = t * (T)1e-3 + (T)k;
        b[k] = (T)1 + t * (T)1e-6;
        y[k] = (T)0.9999 - t * (T)1e-7 * (T)k;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < CAL_CHAINS; ++k) {
            const T m = SGPR ? s : y[k];
            a[k] = fma(b[k], m, a[k]);
            b[k] = fma(a[k], m, b[k]);
        }
    }
    T r = 0;
#pragma unroll
    for (int k = 0; k < CAL_CHAINS; ++k) r += a[k] + b[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = (float)r;
}

}  // namespace tct

using namespace tct;

extern "C" int tct_calib_valu(int mode, int blocks, int iters, float *out, void *stream) {
    TCT_CHECK_ARG(mode >= 0 && mode <= 3, "calib_valu: mode %d", mode);
    TCT_CHECK_ARG(blocks >= 1 && iters >= 1 && out, "calib_valu: bad arguments");
    hipStream_t st = as_stream(stream);
    switch (mode) {
    case 0: hipLaunchKernelGGL((calib_fma_kernel<float, false>), dim3(blocks), dim3(256), 0, st, 0.999f, iters, out); break;
    case 1: hipLaunchKernelGGL((calib_fma_kernel<float, true>), dim3(blocks), dim3(256), 0, st, 0.999f, iters, out); break;
    case 2: hipLaunchKernelGGL((calib_fma_kernel<double, false>), dim3(blocks), dim3(256), 0, st, 0.999, iters, out); break;
    default: hipLaunchKernelGGL((calib_fma_kernel<double, true>), dim3(blocks), dim3(256), 0, st, 0.999, iters, out); break;
    }
    TCT_CHECK_LAUNCH("calib_valu");
    return TCT_OK;
}
