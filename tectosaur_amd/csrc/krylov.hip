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
// timing, and tct_cgs2 uses no library workspace: solves on different
// streams share nothing (tct_cgs2_noalloc instead uses one static array per
// device, below).  HBM-bound: each pass reads V[:m] once.
#include <mutex>

#include "tct_common.h"

namespace tct {

constexpr int KB = 256;  // threads per block

constexpr int CB = 8;  // columns per pass of the dot products

// The block sums of CB per-thread values, in a fixed order (xor butterfly in
// each wavefront, then the wavefronts in order): column i0 + k of block b
// goes to partial[(i0 + k) * nb + b] for i0 + k < cols.  Partial sums are
// stored column by column, [cols][nb], so a column's nb partials are
// contiguous for the column reduction.  red is double-buffered by pass
// parity: a pass writes red[(i0 / CB) & 1] before its barrier and reads it
// after, and the next writer of that half comes after the next pass's
// barrier, so one barrier per pass suffices.
__device__ __forceinline__ void store_block_sums(double (&s)[CB], double (*red)[KB / 64][CB],
                                                 int i0, int cols, int nb,
                                                 double *__restrict__ partial) {
    double(*r2)[CB] = red[(i0 / CB) & 1];
#pragma unroll
    for (int k = 0; k < CB; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
    if (threadIdx.x % 64 == 0)
#pragma unroll
        for (int k = 0; k < CB; ++k) r2[threadIdx.x / 64][k] = s[k];
    __syncthreads();
    if (threadIdx.x < CB && i0 + (int)threadIdx.x < cols) {
        double r = 0.0;
        for (int wv = 0; wv < KB / 64; ++wv) r += r2[wv][threadIdx.x];
        partial[(size_t)(i0 + threadIdx.x) * nb + blockIdx.x] = r;
    }
}

// A block's points [t0, t1) are taken in rounds of PU * KB: in a round,
// thread x holds points tb + x + k KB, k < PU, so every load of a round is
// issued before the first is used.  A thread's points come in the order
// t0 + x, t0 + x + KB, ... whatever PU is, so the per-thread sums (and the
// results) do not depend on PU.
constexpr int PU = 2;

// The CB dot products of one column pass over the block's points: s[k] +=
// X_k[t] w[t] for column i0 + k < cols, where X_k = V[i0 + k] for i0 + k <
// m and X_k = w for i0 + k == m (the norm column).
__device__ __forceinline__ void dot_pass(double (&s)[CB], int i0, int cols, int m, int64_t ld,
                                         int64_t t0, int64_t t1,
                                         const double *__restrict__ V,
                                         const double *__restrict__ w) {
#pragma unroll
    for (int k = 0; k < CB; ++k) s[k] = 0.0;
    for (int64_t tb = t0; tb < t1; tb += PU * KB) {
        double wt[PU];
        bool ok[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int64_t t = tb + u * KB + threadIdx.x;
            ok[u] = t < t1;
            wt[u] = ok[u] ? w[t] : 0.0;
        }
        double x[PU][CB];
#pragma unroll
        for (int u = 0; u < PU; ++u)
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const int i = i0 + k;
                const int64_t t = tb + u * KB + threadIdx.x;
                x[u][k] = (ok[u] && i < m) ? V[(size_t)i * ld + t] : wt[u];
            }
#pragma unroll
        for (int u = 0; u < PU; ++u)
#pragma unroll
            for (int k = 0; k < CB; ++k)
                if (ok[u] && i0 + k < cols) s[k] = fma(x[u][k], wt[u], s[k]);
    }
}

// partial[i][b] = sum over block b's points of V[i][t] w[t], i < m, and
// partial[m][b] = sum w[t]^2 when with_norm; CB columns per pass over the
// block's points (one barrier pair per pass).
__global__ __launch_bounds__(KB) void multi_dot_kernel(int m, int64_t n, int64_t ld,
                                                       const double *__restrict__ V,
                                                       const double *__restrict__ w,
                                                       int with_norm,
                                                       double *__restrict__ partial) {
    __shared__ double red[2][KB / 64][CB];
    const int nb = gridDim.x;
    const int cols = m + (with_norm ? 1 : 0);
    // this block's contiguous range of points
    const int64_t per = (n + nb - 1) / nb;
    const int64_t t0 = (int64_t)blockIdx.x * per;
    const int64_t t1 = t0 + per < n ? t0 + per : n;
    for (int i0 = 0; i0 < cols; i0 += CB) {
        double s[CB];
        dot_pass(s, i0, cols, m, ld, t0, t1, V, w);
        store_block_sums(s, red, i0, cols, nb, partial);
    }
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

// Column sums of the per-block partials, one wavefront per column: c[i] =
// sum_b partial[i][b] (each lane takes every 64th block in order, then an
// xor butterfly: a fixed order), hout[i] = c[i] (+ add[i]).  Used when there
// are many blocks (below).
__global__ __launch_bounds__(KB) void reduce_cols_kernel(int m, int nb,
                                                         const double *__restrict__ partial,
                                                         const double *__restrict__ add,
                                                         double *__restrict__ c,
                                                         double *__restrict__ hout) {
    const int i = blockIdx.x * (KB / 64) + (int)threadIdx.x / 64;
    const int lane = (int)threadIdx.x % 64;
    if (i >= m) return;   // the whole wavefront
    const double *p = partial + (size_t)i * nb;
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += p[b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        c[i] = s;
        if (hout) hout[i] = add ? s + add[i] : s;
    }
}

// The reduce, the update and the next sums in one launch, on the partition
// of multi_dot_kernel (gridDim.x == nb blocks of contiguous points): the
// column sums c[i] of partial_in[i][0..nb) in block order -- summed by every
// block itself, block 0 writing c (+ add) to hout, or, with many blocks,
// read from c_in where reduce_cols_kernel left them (without nb re-reads of
// the m x nb partials); w[t] -= sum_i V[i][t] c[i] on the block's points;
// then partial_out[i][b] = sum over those points of V[i][t] w[t] (norm_only:
// the one column sum w[t]^2), as multi_dot_kernel forms them.  Each w[t] is
// updated by one sum in a fixed order, so the values do not depend on the
// launch shape; each thread reads back only the w[t] it wrote.  partial_in
// and partial_out must not overlap.
__global__ __launch_bounds__(KB) void reduce_update_dot_kernel(
    int m, int64_t n, int64_t ld, const double *__restrict__ V,
    const double *__restrict__ partial_in, const double *__restrict__ c_in,
    const double *__restrict__ add, double *__restrict__ hout, double *__restrict__ w,
    int norm_only, double *__restrict__ partial_out) {
    extern __shared__ double c[];
    __shared__ double red[2][KB / 64][CB];
    const int nb = gridDim.x;
    for (int i = threadIdx.x; i < m; i += KB) {
        if (c_in) {
            c[i] = c_in[i];
        } else {
            double s = 0.0;
            for (int b = 0; b < nb; ++b) s += partial_in[(size_t)i * nb + b];
            c[i] = s;
            if (blockIdx.x == 0 && hout) hout[i] = add ? s + add[i] : s;
        }
    }
    __syncthreads();
    const int64_t per = (n + nb - 1) / nb;
    const int64_t t0 = (int64_t)blockIdx.x * per;
    const int64_t t1 = t0 + per < n ? t0 + per : n;
    // w[t] -= sum_i V[i][t] c[i], i ascending; IU rows' loads of a round's
    // PU points in flight at once
    constexpr int IU = 4;
    for (int64_t tb = t0; tb < t1; tb += PU * KB) {
        double acc[PU];
        bool ok[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            ok[u] = tb + u * KB + threadIdx.x < t1;
            acc[u] = 0.0;
        }
        const double *vp = V + tb + threadIdx.x;
        int i = 0;
        for (; i + IU <= m; i += IU) {
            double x[IU][PU];
#pragma unroll
            for (int r = 0; r < IU; ++r)
#pragma unroll
                for (int u = 0; u < PU; ++u)
                    x[r][u] = ok[u] ? vp[(size_t)(i + r) * ld + u * KB] : 0.0;
#pragma unroll
            for (int r = 0; r < IU; ++r)
#pragma unroll
                for (int u = 0; u < PU; ++u) acc[u] = fma(x[r][u], c[i + r], acc[u]);
        }
        for (; i < m; ++i)
#pragma unroll
            for (int u = 0; u < PU; ++u)
                acc[u] = fma(ok[u] ? vp[(size_t)i * ld + u * KB] : 0.0, c[i], acc[u]);
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            const int64_t t = tb + u * KB + threadIdx.x;
            if (ok[u]) w[t] = w[t] - acc[u];
        }
    }
    // each thread reads back only the w[t] it wrote
    const int cols = norm_only ? 1 : m;
    for (int i0 = 0; i0 < cols; i0 += CB) {
        double s[CB];
        dot_pass(s, i0, cols, norm_only ? 0 : m, ld, t0, t1, V, w);
        store_block_sums(s, red, i0, cols, nb, partial_out);
    }
}

// |w| from the norm partials (every block, same order), written to *hnorm
// by block 0, and vnext (vnext2) = w / |w| when given.
__global__ __launch_bounds__(KB) void norm_normalize_kernel(int64_t n,
                                                            const double *__restrict__ w,
                                                            const double *__restrict__ partial,
                                                            int nb, double *__restrict__ hnorm,
                                                            double *__restrict__ vnext,
                                                            double *__restrict__ vnext2) {
    __shared__ double nrm;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partial[b];
        s = sqrt(s);
        nrm = s;
        if (blockIdx.x == 0) *hnorm = s;
    }
    __syncthreads();
    if (!vnext) return;
    const double d = nrm;
    for (int64_t t = (int64_t)blockIdx.x * KB + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * KB) {
        const double v = w[t] / d;
        vnext[t] = v;
        if (vnext2) vnext2[t] = v;
    }
}

// Up to GATHER_MAX device segments concatenated into one buffer: block k
// copies segment k (kernel arguments by value, no table upload).
constexpr int GATHER_MAX = 32;
struct GatherSegs {
    const double *src[GATHER_MAX];
    int off[GATHER_MAX + 1];
};

__global__ __launch_bounds__(KB) void gather_segs_kernel(GatherSegs segs,
                                                         double *__restrict__ out) {
    const int k = blockIdx.x;
    const double *src = segs.src[k];
    const int o = segs.off[k], len = segs.off[k + 1] - o;
    for (int i = threadIdx.x; i < len; i += KB) out[o + i] = src[i];
}

inline int n_blocks(int64_t n) {
    int64_t nb = (n + 4 * KB - 1) / (4 * KB);
    return (int)(nb < 1 ? 1 : (nb > 1024 ? 1024 : nb));
}

// From this many blocks on, the column sums of the partials are formed once
// by reduce_cols_kernel (two more launches per step) instead of by every
// block of the update (nb^2 m re-reads).
constexpr int NB_SEPARATE_REDUCE = 128;

// The four (or six) launches of one CGS2 step; scratch = p1 [m][nb],
// p2 [m][nb], h1 [m], c [m].
inline void cgs2_launch(int m, int64_t n, int64_t ld, const double *V, double *w, double *h,
                        double *scratch, double *vnext, double *vnext2, hipStream_t st) {
    const int nb = n_blocks(n);
    double *p1 = scratch;                     // [m][nb], later the norm partials
    double *p2 = scratch + (size_t)nb * m;    // [m][nb]
    double *h1 = scratch + 2 * (size_t)nb * m;
    double *cs = h1 + m;                      // column sums (separate reduce)
    const bool sep = nb >= NB_SEPARATE_REDUCE;
    const unsigned gr = (unsigned)ceil_div((int64_t)m, (int64_t)(KB / 64));
    // normalisation launch: enough blocks to cover the points, at most 1024
    const int64_t gpl = ceil_div(n, (int64_t)KB);
    const unsigned gp = (unsigned)(gpl < 1024 ? gpl : 1024);
    const size_t lds = sizeof(double) * (size_t)m;
    // p1 = partial sums of V w
    hipLaunchKernelGGL(multi_dot_kernel, dim3(nb), dim3(KB), 0, st, m, n, ld, V, w, 0, p1);
    // h1 = sum p1; w -= V^T h1; p2 = partial sums of V w
    if (sep)
        hipLaunchKernelGGL(reduce_cols_kernel, dim3(gr), dim3(KB), 0, st, m, nb,
                           (const double *)p1, (const double *)nullptr, cs, h1);
    hipLaunchKernelGGL(reduce_update_dot_kernel, dim3(nb), dim3(KB), lds, st, m, n, ld, V,
                       (const double *)p1, (const double *)(sep ? cs : nullptr),
                       (const double *)nullptr, h1, w, 0, p2);
    // h2 = sum p2, h = h1 + h2; w -= V^T h2; p1 = partial sums of w^2
    if (sep)
        hipLaunchKernelGGL(reduce_cols_kernel, dim3(gr), dim3(KB), 0, st, m, nb,
                           (const double *)p2, (const double *)h1, cs, h);
    hipLaunchKernelGGL(reduce_update_dot_kernel, dim3(nb), dim3(KB), lds, st, m, n, ld, V,
                       (const double *)p2, (const double *)(sep ? cs : nullptr),
                       (const double *)h1, h, w, 1, p1);
    // h[m] = |w|, and the next basis vector
    hipLaunchKernelGGL(norm_normalize_kernel, dim3(vnext ? gp : 1), dim3(KB), 0, st, n, w, p1, nb,
                       h + m, vnext, vnext2);
}

}  // namespace tct

using namespace tct;

extern "C" int tct_krylov_scratch(int m, int64_t n, int64_t *n_doubles) {
    TCT_CHECK_ARG(m >= 1 && n >= 1 && n_doubles, "krylov_scratch: bad arguments");
    // two partial arrays of m columns per block (one read while the other
    // is written), plus h1 [m] and the column sums [m]
    *n_doubles = 2 * (int64_t)n_blocks(n) * m + 2 * m;
    return TCT_OK;
}

extern "C" int tct_cgs2(int m, int64_t n, int64_t ld, const double *V, double *w, double *h,
                        double *scratch, int64_t scratch_doubles, double *vnext, double *vnext2,
                        void *stream) {
    TCT_CHECK_ARG(m >= 1 && n >= 1 && ld >= n && m < (1 << 20), "cgs2: m %d n %lld ld %lld", m,
                  (long long)n, (long long)ld);
    TCT_CHECK_ARG(V && w && h && scratch, "cgs2: null pointer");
    const int nb = n_blocks(n);
    TCT_CHECK_ARG(scratch_doubles >= 2 * (int64_t)nb * m + 2 * m, "cgs2: scratch too small");
    TCT_CHECK_ARG(sizeof(double) * (size_t)m <= 60 * 1024, "cgs2: basis size %d too large for LDS",
                  m);
    cgs2_launch(m, n, ld, V, w, h, scratch, vnext, vnext2, as_stream(stream));
    TCT_CHECK_LAUNCH("cgs2");
    return TCT_OK;
}

// ---- CGS2 without a workspace argument ------------------------------------
// The partial sums live in a file-scope device array instead of a caller
// buffer, for callers that may not allocate (code under graph capture,
// serving loops).  Each device holds its own copy for the life of the
// process; its address is looked up once per device with
// hipGetSymbolAddress and only that address is passed to kernels (the
// array's name is not a device address in host code).  Every kernel writes
// the part of the array it later reads, so nothing needs clearing between
// calls; calls that overlap on one device share the array and race.
constexpr int64_t CGS2_NOALLOC_MAX_DOUBLES = TCT_CGS2_NOALLOC_MAX_DOUBLES;  // 32 MiB

static __device__ __align__(256) double scratch[CGS2_NOALLOC_MAX_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// This device's address of `scratch` (looked up on the first call per device).
static int scratch_address(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return set_error(TCT_EHIP, "cgs2_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_error(TCT_EARG, "cgs2_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(TCT_EHIP, "cgs2_noalloc: hipGetSymbolAddress: %s",
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return TCT_OK;
}

extern "C" int tct_cgs2_noalloc(int m, int64_t n, int64_t ld, const double *V, double *w,
                                double *h, double *vnext, double *vnext2, void *stream) {
    TCT_CHECK_ARG(m >= 1 && n >= 1 && ld >= n && m < (1 << 20),
                  "cgs2_noalloc: m %d n %lld ld %lld", m, (long long)n, (long long)ld);
    TCT_CHECK_ARG(V && w && h, "cgs2_noalloc: null pointer");
    TCT_CHECK_ARG(sizeof(double) * (size_t)m <= 60 * 1024,
                  "cgs2_noalloc: basis size %d too large for LDS", m);
    const int nb = n_blocks(n);
    const int64_t need = 2 * (int64_t)nb * m + 2 * m;
    TCT_CHECK_ARG(need <= CGS2_NOALLOC_MAX_DOUBLES,
                  "cgs2_noalloc: m %d, n %lld need %lld doubles of scratch, more than the "
                  "static %lld",
                  m, (long long)n, (long long)need, (long long)CGS2_NOALLOC_MAX_DOUBLES);
    double *s = nullptr;
    const int rc = scratch_address(&s);
    if (rc != TCT_OK) return rc;
    cgs2_launch(m, n, ld, V, w, h, s, vnext, vnext2, as_stream(stream));
    TCT_CHECK_LAUNCH("cgs2_noalloc");
    return TCT_OK;
}

extern "C" int tct_arnoldi_step(const void *graph_exec, int m, int64_t n, int64_t ld,
                                const double *V, double *w, double *h, double *scratch,
                                int64_t scratch_doubles, double *vnext, double *vnext2,
                                void *stream) {
    TCT_CHECK_ARG(graph_exec, "arnoldi_step: null graph");
    hipError_t e = hipGraphLaunch((hipGraphExec_t)graph_exec, as_stream(stream));
    if (e != hipSuccess)
        return set_error(TCT_EHIP, "arnoldi_step: graph launch: %s", hipGetErrorString(e));
    return tct_cgs2(m, n, ld, V, w, h, scratch, scratch_doubles, vnext, vnext2, stream);
}

extern "C" int tct_tick_gather(int n_streams, void *const *streams, void *const *events,
                               int n_seg, const double *const *src, const int *len,
                               double *staging, int64_t staging_doubles, double *host,
                               void *main_stream, void *done_event) {
    TCT_CHECK_ARG(n_streams >= 0 && n_seg >= 0, "tick_gather: negative counts");
    TCT_CHECK_ARG(n_streams == 0 || (streams && events), "tick_gather: null stream list");
    TCT_CHECK_ARG(n_seg == 0 || (src && len && staging && host), "tick_gather: null pointer");
    int64_t total = 0;
    for (int k = 0; k < n_seg; ++k) {
        TCT_CHECK_ARG(len[k] >= 0 && (len[k] == 0 || src[k]), "tick_gather: bad segment %d", k);
        total += len[k];
    }
    TCT_CHECK_ARG(total <= staging_doubles, "tick_gather: %lld values exceed the staging buffer",
                  (long long)total);
    TCT_CHECK_ARG(total < (1LL << 31), "tick_gather: %lld values (offsets are 32-bit)",
                  (long long)total);
    hipStream_t st = as_stream(main_stream);
    hipError_t e = hipSuccess;
    for (int i = 0; i < n_streams && e == hipSuccess; ++i) {
        e = hipEventRecord((hipEvent_t)events[i], as_stream(streams[i]));
        if (e == hipSuccess) e = hipStreamWaitEvent(st, (hipEvent_t)events[i], 0);
    }
    if (e != hipSuccess)
        return set_error(TCT_EHIP, "tick_gather: stream join: %s", hipGetErrorString(e));
    int64_t off = 0;
    for (int k0 = 0; k0 < n_seg; k0 += GATHER_MAX) {
        const int nk = n_seg - k0 < GATHER_MAX ? n_seg - k0 : GATHER_MAX;
        GatherSegs segs;
        int64_t o = 0;
        for (int k = 0; k < nk; ++k) {
            segs.src[k] = src[k0 + k];
            segs.off[k] = (int)o;
            o += len[k0 + k];
        }
        segs.off[nk] = (int)o;
        if (o > 0)
            hipLaunchKernelGGL(gather_segs_kernel, dim3(nk), dim3(KB), 0, st, segs, staging + off);
        off += o;
    }
    TCT_CHECK_LAUNCH("tick_gather");
    if (total > 0)
        e = hipMemcpyAsync(host, staging, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && done_event) e = hipEventRecord((hipEvent_t)done_event, st);
    if (e != hipSuccess)
        return set_error(TCT_EHIP, "tick_gather: copy: %s", hipGetErrorString(e));
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
