// Krylov-basis kernels of the device GMRES: classical Gram-Schmidt repeated
// twice (CGS2) against the basis V [m][n] (row-major, row stride ld), and
// the basis combination d = V^T y.
//
// Reference: the Arnoldi step of scipy.sparse.linalg.gmres that
// tectosaur/simple_solver.py:14-34 calls (parity unpinned against SciPy;
// the solves are checked by their FP64 residual).
//
// Every reduction has a fixed order (per-block partials over fixed point
// ranges, then one pass over the blocks), so results do not depend on
// timing, and no library workspace is involved: solves on different
// streams share nothing.  HBM-bound: each pass reads V[:m] once.
#include "tct_common.h"

namespace tct {

constexpr int KB = 256;  // threads per block

// partial[b][i] = sum over block b's points of V[i][t] w[t], i < m, and
// partial[b][m] = sum w[t]^2 when with_norm.  Grid-stride over t.
__global__ __launch_bounds__(KB) void multi_dot_kernel(int m, int64_t n, int64_t ld,
                                                       const double *__restrict__ V,
                                                       const double *__restrict__ w,
                                                       int with_norm,
                                                       double *__restrict__ partial) {
    __shared__ double red[KB / 64];
    const int nb = gridDim.x;
    const int cols = m + (with_norm ? 1 : 0);
    // this block's contiguous range of points
    const int64_t per = (n + nb - 1) / nb;
    const int64_t t0 = (int64_t)blockIdx.x * per;
    const int64_t t1 = t0 + per < n ? t0 + per : n;
    for (int i = 0; i < cols; ++i) {
        const double *vi = i < m ? V + (size_t)i * ld : w;
        double s = 0.0;
        for (int64_t t = t0 + threadIdx.x; t < t1; t += KB) s += vi[t] * w[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int k = 0; k < KB / 64; ++k) r += red[k];
            partial[(size_t)blockIdx.x * cols + i] = r;
        }
        __syncthreads();
    }
}

// h[i] (+)= sum_b partial[b][i] for i < cols; with sqrt_last the last
// column is a sum of squares and its square root is written.
__global__ void reduce_cols_kernel(int cols, int nb, const double *__restrict__ partial,
                                   const double *__restrict__ add, int sqrt_last,
                                   double *__restrict__ h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[(size_t)b * cols + i];
    if (sqrt_last && i == cols - 1)
        s = sqrt(s);
    else if (add)
        s += add[i];
    h[i] = s;
}

// out[t] = beta out[t] + alpha sum_i V[i][t] c[i]   (one thread per point)
__global__ __launch_bounds__(KB) void combine_kernel(int m, int64_t n, int64_t ld,
                                                     const double *__restrict__ V,
                                                     const double *__restrict__ c, double alpha,
                                                     double beta, double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (t >= n) return;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += V[(size_t)i * ld + t] * c[i];
    out[t] = (beta == 0.0 ? 0.0 : beta * out[t]) + alpha * s;
}

// vnext[t] = w[t] / *norm (and vnext2 likewise): the next basis vector,
// formed on the device from the norm the step just computed.
__global__ __launch_bounds__(KB) void normalize_kernel(int64_t n, const double *__restrict__ w,
                                                       const double *__restrict__ norm,
                                                       double *__restrict__ vnext,
                                                       double *__restrict__ vnext2) {
    const int64_t t = (int64_t)blockIdx.x * KB + threadIdx.x;
    if (t >= n) return;
    const double v = w[t] / *norm;
    vnext[t] = v;
    if (vnext2) vnext2[t] = v;
}

inline int n_blocks(int64_t n) {
    int64_t nb = (n + 4 * KB - 1) / (4 * KB);
    return (int)(nb < 1 ? 1 : (nb > 1024 ? 1024 : nb));
}

}  // namespace tct

using namespace tct;

extern "C" int tct_krylov_scratch(int m, int64_t n, int64_t *n_doubles) {
    TCT_CHECK_ARG(m >= 1 && n >= 1 && n_doubles, "krylov_scratch: bad arguments");
    // partials for m + 1 columns, plus h1 [m]
    *n_doubles = (int64_t)n_blocks(n) * (m + 1) + m;
    return TCT_OK;
}

extern "C" int tct_cgs2(int m, int64_t n, int64_t ld, const double *V, double *w, double *h,
                        double *scratch, int64_t scratch_doubles, double *vnext, double *vnext2,
                        void *stream) {
    TCT_CHECK_ARG(m >= 1 && n >= 1 && ld >= n && m < (1 << 20), "cgs2: m %d n %lld ld %lld", m,
                  (long long)n, (long long)ld);
    TCT_CHECK_ARG(V && w && h && scratch, "cgs2: null pointer");
    const int nb = n_blocks(n);
    TCT_CHECK_ARG(scratch_doubles >= (int64_t)nb * (m + 1) + m, "cgs2: scratch too small");
    hipStream_t st = as_stream(stream);
    double *partial = scratch;
    double *h1 = scratch + (size_t)nb * (m + 1);
    const unsigned gp = (unsigned)ceil_div(n, KB);
    const unsigned gr = (unsigned)ceil_div(m + 1, 64);
    // pass 1: h1 = V w; w -= V^T h1
    hipLaunchKernelGGL(multi_dot_kernel, dim3(nb), dim3(KB), 0, st, m, n, ld, V, w, 0, partial);
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(gr), dim3(64), 0, st, m, nb, partial,
                       (const double *)nullptr, 0, h1);
    hipLaunchKernelGGL(combine_kernel, dim3(gp), dim3(KB), 0, st, m, n, ld, V, h1, -1.0, 1.0, w);
    // pass 2: h2 = V w; w -= V^T h2; h = h1 + h2
    hipLaunchKernelGGL(multi_dot_kernel, dim3(nb), dim3(KB), 0, st, m, n, ld, V, w, 0, partial);
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(gr), dim3(64), 0, st, m, nb, partial,
                       (const double *)nullptr, 0, h);
    hipLaunchKernelGGL(combine_kernel, dim3(gp), dim3(KB), 0, st, m, n, ld, V, h, -1.0, 1.0, w);
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(gr), dim3(64), 0, st, m, 1, h, h1, 0, h);
    // h[m] = |w|
    hipLaunchKernelGGL(multi_dot_kernel, dim3(nb), dim3(KB), 0, st, 0, n, ld, V, w, 1, partial);
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(64), 0, st, 1, nb, partial,
                       (const double *)nullptr, 1, h + m);
    if (vnext)
        hipLaunchKernelGGL(normalize_kernel, dim3(gp), dim3(KB), 0, st, n, w, h + m, vnext,
                           vnext2);
    TCT_CHECK_LAUNCH("cgs2");
    return TCT_OK;
}

extern "C" int tct_krylov_combine(int m, int64_t n, int64_t ld, const double *V, const double *c,
                                  double alpha, double beta, double *out, void *stream) {
    TCT_CHECK_ARG(m >= 0 && n >= 1 && ld >= n, "krylov_combine: m %d n %lld", m, (long long)n);
    TCT_CHECK_ARG(out && (m == 0 || (V && c)), "krylov_combine: null pointer");
    hipLaunchKernelGGL(combine_kernel, dim3((unsigned)ceil_div(n, KB)), dim3(KB), 0,
                       as_stream(stream), m, n, ld, V, c, alpha, beta, out);
    TCT_CHECK_LAUNCH("krylov_combine");
    return TCT_OK;
}
