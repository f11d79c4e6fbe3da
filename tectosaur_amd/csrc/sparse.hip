// Block-sparse and CSR matrix-vector products, FP64, gfx950.
//
// bsr9: the nearfield operator applied every GMRES iteration, reference
// bsrmv9 / bcoomv9 (tectosaur/util/fast_sparse.cpp:71-163).  The reference
// runs it on the host with an OpenMP atomic per output element; here the
// blocks are row-sorted (BSR) so every output row has one writer and no
// atomics are needed.
//
// Layout: one wavefront per block row.  63 of its lanes are (b, i) pairs,
// b < 7 a block of the current group of 7 consecutive blocks and i < 9 a
// row of that block; a lane reads its block row (9 contiguous doubles) and
// the 9 doubles of x it multiplies, so each step of the wave streams 7
// whole blocks (4.5 KB contiguous).  The 7 partial sums per output row are
// added in fixed order through LDS at the end.  HBM-bound: 648 B of block
// data per block against 162 flops.
#include "tct_common.h"

namespace tct {

template <typename TD>
__global__ __launch_bounds__(256) void bsr9_kernel(int n_block_rows,
                                                   const int32_t *__restrict__ indptr,
                                                   const int32_t *__restrict__ indices,
                                                   const TD *__restrict__ data,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ y, int accumulate) {
    constexpr int NB = 7;
    __shared__ double part[4][NB * 9];
    __shared__ int scol[4][64];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    const bool live = row < n_block_rows;
    const int b = lane / 9, i = lane % 9;
    double s = 0.0;
    if (live) {
        const int k0 = indptr[row], k1 = indptr[row + 1];
        // the column indices of the next 63 blocks (9 steps of NB) come in
        // one coalesced load through the wave's LDS slot, so no block waits
        // on its own index load before its data and x loads can issue
        for (int kb = k0; kb < k1; kb += 9 * NB) {
            scol[wave][lane] = kb + lane < k1 ? indices[kb + lane] : 0;
            __builtin_amdgcn_wave_barrier();
            const int ke = kb + 9 * NB < k1 ? kb + 9 * NB : k1;
            if (lane < NB * 9) {
#pragma unroll 2
                for (int k = kb + b; k < ke; k += NB) {
                    const TD *A = data + (size_t)k * 81 + i * 9;
                    const double *xc = x + (size_t)scol[wave][k - kb] * 9;
#pragma unroll
                    for (int j = 0; j < 9; ++j) s += (double)A[j] * xc[j];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (lane < NB * 9) part[wave][lane] = s;
    __syncthreads();
    if (live && lane < 9) {
        double t = 0.0;
#pragma unroll
        for (int bb = 0; bb < NB; ++bb) t += part[wave][bb * 9 + lane];
        double *yr = y + (size_t)row * 9 + lane;
        *yr = accumulate ? *yr + t : t;
    }
}

// One thread per row.  The row's entries are taken RC at a time: their
// column indices and values load together, then the RC gathered x values,
// then the products are added in entry order (one fused multiply-add each),
// so a row waits on memory twice per RC entries rather than twice per entry.
__global__ __launch_bounds__(256) void csr_kernel(int n_rows, const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ indices,
                                                  const double *__restrict__ data,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y, int accumulate) {
    constexpr int RC = 4;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const int k1 = indptr[r + 1];
    int k = indptr[r];
    double s = 0.0;
    for (; k + RC <= k1; k += RC) {
        int c[RC];
        double a[RC], xv[RC];
#pragma unroll
        for (int j = 0; j < RC; ++j) {
            c[j] = indices[k + j];
            a[j] = data[k + j];
        }
#pragma unroll
        for (int j = 0; j < RC; ++j) xv[j] = x[c[j]];
#pragma unroll
        for (int j = 0; j < RC; ++j) s = fma(a[j], xv[j], s);
    }
    for (; k < k1; ++k) s = fma(data[k], x[indices[k]], s);
    y[r] = accumulate ? y[r] + s : s;
}

}  // namespace tct

using namespace tct;

extern "C" int tct_bsr9_spmv(int64_t n_block_rows, const int32_t *indptr,
                             const int32_t *indices, const double *data, const double *x,
                             double *y, int accumulate, void *stream) {
    TCT_CHECK_ARG(n_block_rows >= 0 && n_block_rows < (1LL << 27), "bsr9: n_block_rows");
    TCT_CHECK_ARG(indptr && x && y, "bsr9: null pointer");
    if (n_block_rows == 0) return TCT_OK;
    hipLaunchKernelGGL(bsr9_kernel<double>, dim3((unsigned)ceil_div(n_block_rows, 4)), dim3(256),
                       0, as_stream(stream), (int)n_block_rows, indptr, indices, data, x, y,
                       accumulate);
    TCT_CHECK_LAUNCH("bsr9_spmv");
    return TCT_OK;
}

extern "C" int tct_bsr9_spmv_f32(int64_t n_block_rows, const int32_t *indptr,
                                 const int32_t *indices, const float *data, const double *x,
                                 double *y, int accumulate, void *stream) {
    TCT_CHECK_ARG(n_block_rows >= 0 && n_block_rows < (1LL << 27), "bsr9_f32: n_block_rows");
    TCT_CHECK_ARG(indptr && x && y, "bsr9_f32: null pointer");
    if (n_block_rows == 0) return TCT_OK;
    hipLaunchKernelGGL(bsr9_kernel<float>, dim3((unsigned)ceil_div(n_block_rows, 4)), dim3(256),
                       0, as_stream(stream), (int)n_block_rows, indptr, indices, data, x, y,
                       accumulate);
    TCT_CHECK_LAUNCH("bsr9_spmv_f32");
    return TCT_OK;
}

extern "C" int tct_csr_spmv(int64_t n_rows, const int32_t *indptr, const int32_t *indices,
                            const double *data, const double *x, double *y, int accumulate,
                            void *stream) {
    TCT_CHECK_ARG(n_rows >= 0 && n_rows < (1LL << 31), "csr: n_rows");
    TCT_CHECK_ARG(indptr && x && y, "csr: null pointer");
    if (n_rows == 0) return TCT_OK;
    hipLaunchKernelGGL(csr_kernel, dim3((unsigned)ceil_div(n_rows, 256)), dim3(256), 0,
                       as_stream(stream), (int)n_rows, indptr, indices, data, x, y, accumulate);
    TCT_CHECK_LAUNCH("csr_spmv");
    return TCT_OK;
}
