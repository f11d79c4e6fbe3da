// Treecode evaluation kernels for TCT_RT (orders 2-6, FP32 and FP64).
#include "far_tree.h"

namespace tct {

template <>
int tree_eval_kernel_entry<TCT_RT>(int order, int prec, dim3 grid, hipStream_t st,
                                int n_obs_pad, const void *obs, const double *obs_org,
                                const void *rec, const double *src_org,
                                const double *node_cr, const void *moments,
                                const int *seg, const int *list,
                                const double *params, double eps2, double *partial) {
    return eval_order<TCT_RT>(order, prec, grid, st, n_obs_pad, obs, obs_org, rec, src_org,
                          node_cr, moments, seg, list, params, eps2, partial);
}

}  // namespace tct
