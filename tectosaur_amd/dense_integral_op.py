"""Dense integral operators: ``farfield_tris`` and
``RegularizedDenseIntegralOp``, drop-ins for
tectosaur/ops/dense_integral_op.py:11-84.

``farfield_tris(kernel, params, pts, obs_tris, src_tris, n_q, float_type)``
returns the dense pair matrix [n_obs, 3, 3, n_src, 3, 3] of every
(observation triangle, source triangle) pair integrated with the product
Gauss rule gauss4d_tri(n_q) -- the reference's ``farfield_tris`` kernel
(tectosaur/assemble.cl:105-138), one thread per pair there.  Here the pairs
go through the FP64 pair integrator (pairs.hip) in bounded batches, and
each batch's [pairs, 81] blocks are permuted into the dense layout on the
device.  Pairs whose rule points coincide (a triangle paired with itself)
give inf / NaN there, as in the reference.

``RegularizedDenseIntegralOp`` is the reference's dense operator: the
uncorrected nearfield blocks where they are nonzero, the farfield rule
everywhere else, non-finite entries zeroed (dense_integral_op.py:55-84;
the reference zeroes NaN, which is what the far rule of a triangle paired
with itself gives in FP64; an infinity there, which the device's operation
order can give instead, is zeroed the same way).  It is meant for small
problems (the matrix is (9 n)^2 doubles).
"""
import numpy as np
import torch

from . import _native as N
from .combinators import DenseOp
from .nearfield_op import RegularizedNearfieldIntegralOp
from .pairs_integrator import PairsIntegrator

# (obs, src) pairs per integrator launch: bounds the [pairs, 81] buffer
DENSE_CALL_PAIRS = 1 << 18


def farfield_tris_dev(kernel, params, pts, obs_tris, src_tris, n_q):
    """Device tensor [n_obs * 9, n_src * 9] (FP64) of the farfield pair
    matrix, rows (obs, b_obs, d_obs) and columns (src, b_src, d_src)."""
    obs_tris = np.asarray(obs_tris, dtype=np.int64).reshape((-1, 3))
    src_tris = np.asarray(src_tris, dtype=np.int64).reshape((-1, 3))
    n_obs, n_src = obs_tris.shape[0], src_tris.shape[0]
    all_tris = np.concatenate([obs_tris, src_tris], axis=0)
    pi = PairsIntegrator(kernel, params, np.float64, n_q, n_q, pts, all_tris)
    dev = pi.device
    out = torch.empty((n_obs, 9, n_src, 9), dtype=torch.float64, device=dev)
    src_idx = torch.arange(n_obs, n_obs + n_src, dtype=torch.int32, device=dev)
    rows_per_call = max(1, DENSE_CALL_PAIRS // max(n_src, 1))
    for r0 in range(0, n_obs, rows_per_call):
        r1 = min(n_obs, r0 + rows_per_call)
        obs_idx = torch.arange(r0, r1, dtype=torch.int32, device=dev)
        pairs = torch.stack([obs_idx.repeat_interleave(n_src), src_idx.repeat(r1 - r0)],
                            dim=1).contiguous()
        blk = pi.pairs_quad_dev(pi.far_q, pairs, False)
        # [obs, src, (b_obs d_obs), (b_src d_src)] -> [obs, (b_obs d_obs), src, (b_src d_src)]
        out[r0:r1] = blk.reshape((r1 - r0, n_src, 9, 9)).permute(0, 2, 1, 3)
    return out.reshape((n_obs * 9, n_src * 9))


def farfield_tris(kernel, params, pts, obs_tris, src_tris, n_q, float_type):
    """tectosaur/ops/dense_integral_op.py:11-13: [n_obs, 3, 3, n_src, 3, 3]
    in ``float_type`` (computed in FP64)."""
    n_obs = np.asarray(obs_tris).reshape((-1, 3)).shape[0]
    n_src = np.asarray(src_tris).reshape((-1, 3)).shape[0]
    m = farfield_tris_dev(kernel, params, pts, obs_tris, src_tris, n_q)
    return m.cpu().numpy().reshape((n_obs, 3, 3, n_src, 3, 3)).astype(float_type)


class RegularizedDenseIntegralOp(DenseOp):
    def __init__(self, nq_coincident, nq_edge_adj, nq_vert_adjacent, nq_far, nq_near,
                 near_threshold, K_near_name, K_far_name, params, pts, tris, float_type,
                 obs_subset=None, src_subset=None):
        tris = np.asarray(tris)
        if obs_subset is None:
            obs_subset = np.arange(tris.shape[0])
        if src_subset is None:
            src_subset = np.arange(tris.shape[0])
        obs_subset = np.asarray(obs_subset, dtype=np.int64)
        src_subset = np.asarray(src_subset, dtype=np.int64)
        self.float_type = float_type
        near = RegularizedNearfieldIntegralOp(
            pts, tris, obs_subset, src_subset, nq_coincident, nq_edge_adj, nq_vert_adjacent,
            nq_far, nq_near, near_threshold, K_near_name, K_far_name, params, float_type)
        near_dense = near.mat_no_correction.to_dense_dev()
        far = farfield_tris_dev(K_far_name, params, pts, tris[obs_subset], tris[src_subset],
                                nq_far)
        m = torch.where(near_dense.abs() > 0, near_dense, far)
        m = torch.nan_to_num(m, nan=0.0, posinf=0.0, neginf=0.0)
        super().__init__(m.cpu().numpy())
        self._dev = m
