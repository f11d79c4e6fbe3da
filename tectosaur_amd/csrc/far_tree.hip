// Treecode farfield: tile moments, upward pass and the C ABI.  The kernel
// formulation and layout are described in far_tree.h.
#include "far_tree.h"

namespace tct {

// out [NO][n_obs_pad] = sum over each tile's segments [seg_off[t],
// seg_off[t + 1]) of seg_partial, in segment order.  grid = n_obs_tiles,
// one thread per (output q, point).  The loads of RS segments are issued
// before their values are added (in segment order), so a tile with many
// segments waits on memory once per RS segments, not once per segment.
__global__ __launch_bounds__(576) void tree_reduce_kernel(int no, int n_obs_pad,
                                                          const int *__restrict__ seg_off,
                                                          const double *__restrict__ seg_partial,
                                                          double *__restrict__ out) {
    constexpr int RS = 8;
    const int t = blockIdx.x;
    const int q = threadIdx.x / G, lane = threadIdx.x % G;
    if (q >= no) return;
    const int s0 = seg_off[t], s1 = seg_off[t + 1];
    const size_t stride = (size_t)no * G;
    const double *p = seg_partial + (size_t)q * G + lane;
    double v = 0.0;
    int s = s0;
    for (; s + RS <= s1; s += RS) {
        double x[RS];
#pragma unroll
        for (int k = 0; k < RS; ++k) x[k] = p[(size_t)(s + k) * stride];
#pragma unroll
        for (int k = 0; k < RS; ++k) v += x[k];
    }
    for (; s < s1; ++s) v += p[(size_t)s * stride];
    out[(size_t)q * n_obs_pad + t * G + lane] = v;
}

}  // namespace tct

using namespace tct;

static int tree_kernel_ok(int kernel) {
    return kernel == TCT_U || kernel == TCT_RT || kernel == TCT_RA || kernel == TCT_RH;
}

extern "C" int tct_tree_p2m(int kernel, int prec, int order, int64_t n_src, int64_t n_tiles,
                            const void *rec, const double *src_org, const double *node_cr,
                            void *moments, void *stream) {
    TCT_CHECK_ARG(tree_kernel_ok(kernel), "tree_p2m: kernel %d has no treecode form", kernel);
    TCT_CHECK_ARG(prec == TCT_F32 || prec == TCT_F64, "tree_p2m: precision %d", prec);
    TCT_CHECK_ARG(order >= 0 && order <= TCT_TREE_MAX_ORDER, "tree_p2m: order %d", order);
    TCT_CHECK_ARG(n_tiles > 0 && n_tiles < (1LL << 31) && n_src > (n_tiles - 1) * G &&
                      n_src <= n_tiles * G,
                  "tree_p2m: n_src %lld does not fill n_tiles %lld", (long long)n_src,
                  (long long)n_tiles);
    TCT_CHECK_ARG(rec && node_cr && moments && (prec == TCT_F64 || src_org),
                  "tree_p2m: null pointer");
    hipStream_t st = as_stream(stream);
    const dim3 grid((unsigned)n_tiles), block(256);
#define TCT_P2M(KID)                                                                          \
    if (prec == TCT_F32)                                                                      \
        hipLaunchKernelGGL((tree_p2m_kernel<KID, float, false>), grid, block, 0, st, order,   \
                           n_src, (const float *)rec, src_org, node_cr, (float *)moments);    \
    else                                                                                      \
        hipLaunchKernelGGL((tree_p2m_kernel<KID, double, true>), grid, block, 0, st, order,   \
                           n_src, (const double *)rec, src_org, node_cr, (double *)moments)
    switch (kernel) {
    case TCT_U: TCT_P2M(TCT_U); break;
    case TCT_RT: TCT_P2M(TCT_RT); break;
    case TCT_RA: TCT_P2M(TCT_RA); break;
    default: TCT_P2M(TCT_RH); break;
    }
#undef TCT_P2M
    TCT_CHECK_LAUNCH("tree_p2m");
    return TCT_OK;
}

extern "C" int tct_tree_m2m(int kernel, int prec, int order, int p_off, int n_par, int c_off,
                            int n_child, const double *node_cr, void *moments, void *stream) {
    TCT_CHECK_ARG(tree_kernel_ok(kernel), "tree_m2m: kernel %d has no treecode form", kernel);
    TCT_CHECK_ARG(prec == TCT_F32 || prec == TCT_F64, "tree_m2m: precision %d", prec);
    TCT_CHECK_ARG(order >= 0 && order <= TCT_TREE_MAX_ORDER, "tree_m2m: order %d", order);
    TCT_CHECK_ARG(n_par > 0 && n_child > 0 && n_child <= 2 * n_par && n_child > 2 * (n_par - 1) &&
                      p_off >= 0 && c_off >= 0 && c_off + n_child <= p_off,
                  "tree_m2m: bad level p_off=%d n_par=%d c_off=%d n_child=%d", p_off, n_par,
                  c_off, n_child);
    TCT_CHECK_ARG(node_cr && moments, "tree_m2m: null pointer");
    hipStream_t st = as_stream(stream);
    const dim3 grid((unsigned)n_par), block(256);
#define TCT_M2M(KID)                                                                           \
    if (prec == TCT_F32)                                                                       \
        hipLaunchKernelGGL((tree_m2m_kernel<KID, float>), grid, block, 0, st, order, p_off,    \
                           n_par, c_off, n_child, node_cr, (float *)moments);                  \
    else                                                                                       \
        hipLaunchKernelGGL((tree_m2m_kernel<KID, double>), grid, block, 0, st, order, p_off,   \
                           n_par, c_off, n_child, node_cr, (double *)moments)
    switch (kernel) {
    case TCT_U: TCT_M2M(TCT_U); break;
    case TCT_RT: TCT_M2M(TCT_RT); break;
    case TCT_RA: TCT_M2M(TCT_RA); break;
    default: TCT_M2M(TCT_RH); break;
    }
#undef TCT_M2M
    TCT_CHECK_LAUNCH("tree_m2m");
    return TCT_OK;
}

extern "C" int tct_tree_eval(int kernel, int prec, int order, int n_obs_pad, const void *obs,
                             const double *obs_org, const void *rec, const double *src_org,
                             const double *node_cr, const void *moments, int n_seg,
                             const int *seg, const int *list, const double *params_host,
                             double eps2, double *seg_partial, void *stream) {
    TCT_CHECK_ARG(tree_kernel_ok(kernel), "tree_eval: kernel %d has no treecode form", kernel);
    TCT_CHECK_ARG(prec == TCT_F32 || prec == TCT_F64, "tree_eval: precision %d", prec);
    TCT_CHECK_ARG(n_obs_pad > 0 && n_obs_pad % G == 0, "tree_eval: n_obs_pad %d", n_obs_pad);
    TCT_CHECK_ARG(n_seg > 0, "tree_eval: n_seg %d", n_seg);
    TCT_CHECK_ARG(obs && rec && node_cr && moments && seg && list && params_host &&
                      seg_partial && (prec == TCT_F64 || (obs_org && src_org)),
                  "tree_eval: null pointer");
    hipStream_t st = as_stream(stream);
    const dim3 grid((unsigned)n_seg);
    int rc;
#define TCT_EK(KID)                                                                           \
    rc = tree_eval_kernel_entry<KID>(order, prec, grid, st, n_obs_pad, obs, obs_org, rec,     \
                                     src_org, node_cr, moments, seg, list, params_host, eps2, \
                                     seg_partial)
    switch (kernel) {
    case TCT_U: TCT_EK(TCT_U); break;
    case TCT_RT: TCT_EK(TCT_RT); break;
    case TCT_RA: TCT_EK(TCT_RA); break;
    default: TCT_EK(TCT_RH); break;
    }
#undef TCT_EK
    if (rc != TCT_OK) return rc;
    TCT_CHECK_LAUNCH("tree_eval");
    return TCT_OK;
}

extern "C" int tct_tree_reduce(int kernel, int n_obs_tiles, int n_obs_pad, const int *seg_off,
                               const double *seg_partial, double *out_partial, void *stream) {
    TCT_CHECK_ARG(tree_kernel_ok(kernel), "tree_reduce: kernel %d has no treecode form", kernel);
    TCT_CHECK_ARG(n_obs_tiles > 0 && n_obs_pad == n_obs_tiles * G,
                  "tree_reduce: n_obs_pad %d must be 64 * n_obs_tiles (%d)", n_obs_pad,
                  n_obs_tiles);
    TCT_CHECK_ARG(seg_off && seg_partial && out_partial, "tree_reduce: null pointer");
    const int no = (kernel == TCT_RA || kernel == TCT_RH) ? 9 : 3;
    hipLaunchKernelGGL(tree_reduce_kernel, dim3((unsigned)n_obs_tiles), dim3(G * no), 0,
                       as_stream(stream), no, n_obs_pad, seg_off, seg_partial, out_partial);
    TCT_CHECK_LAUNCH("tree_reduce");
    return TCT_OK;
}
