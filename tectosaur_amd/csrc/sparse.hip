// Block-sparse and CSR matrix-vector products, FP64, gfx950.
//
// bsr9: the nearfield operator applied every GMRES iteration, reference
// bsrmv9 / bcoomv9 (tectosaur/util/fast_sparse.cpp:71-163).  The reference
// runs it on the host with an OpenMP atomic per output element; here the
// blocks are row-sorted (BSR) so every output row has one writer and no
// atomics are needed.
//
// Layout: one 16-lane quarter-wave per block row.  Lane l < 9 owns output
// row i = l of the 9x9 block row and walks all blocks of that block row,
// reading the contiguous 9 doubles of row i of each block (72 B) and the 9
// doubles of x it multiplies.  Four block rows per wavefront keep the loads
// of neighbouring rows in flight together.  The kernel is HBM-bound: 648 B
// of block data per block against 162 flops.
#include "tct_common.h"

namespace tct {

__global__ __launch_bounds__(256) void bsr9_kernel(int n_block_rows,
                                                   const int32_t *__restrict__ indptr,
                                                   const int32_t *__restrict__ indices,
                                                   const double *__restrict__ data,
                                                   const double *__restrict__ x,
                                                   double *__restrict__ y, int accumulate) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = gid >> 4;
    const int i = gid & 15;
    if (row >= n_block_rows || i >= 9) return;
    const int k0 = indptr[row], k1 = indptr[row + 1];
    double s = 0.0;
    for (int k = k0; k < k1; ++k) {
        const double *A = data + (size_t)k * 81 + i * 9;
        const double *xc = x + (size_t)indices[k] * 9;
#pragma unroll
        for (int j = 0; j < 9; ++j) s += A[j] * xc[j];
    }
    double *yr = y + (size_t)row * 9 + i;
    *yr = accumulate ? *yr + s : s;
}

__global__ __launch_bounds__(256) void csr_kernel(int n_rows, const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ indices,
                                                  const double *__restrict__ data,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y, int accumulate) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double s = 0.0;
    for (int k = indptr[r]; k < indptr[r + 1]; ++k) s += data[k] * x[indices[k]];
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
    const int64_t threads = n_block_rows * 16;
    hipLaunchKernelGGL(bsr9_kernel, dim3((unsigned)ceil_div(threads, 256)), dim3(256), 0,
                       as_stream(stream), (int)n_block_rows, indptr, indices, data, x, y,
                       accumulate);
    TCT_CHECK_LAUNCH("bsr9_spmv");
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
