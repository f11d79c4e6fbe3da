"""``RegularizedSparseIntegralOp``: the summed matvec GMRES calls, drop-in for
tectosaur/ops/sparse_integral_op.py:10-62.

    y = farfield(v) + nearfield(v)

The reference overlaps a GPU farfield with a host nearfield through asyncio
(:45-55).  Here both halves run on the device on one stream: the farfield
writes y, the nearfield BSR kernel accumulates into it, so a matvec is four
kernel launches with no host round trip (``dot_dev``).  ``dot`` keeps the
host-array API.
"""
import numpy as np
import torch

from .nearfield_op import RegularizedNearfieldIntegralOp
from .sparse_farfield_op import TriToTriDirectFarfieldOp


class RegularizedSparseIntegralOp:
    def __init__(self, nq_coincident, nq_edge_adj, nq_vert_adjacent, nq_far, nq_near,
                 near_threshold, K_near_name, K_far_name, params, pts, tris, float_type,
                 farfield_op_type=TriToTriDirectFarfieldOp, obs_subset=None, src_subset=None,
                 farfield_precisions=None, near_plan=None, far_plan=None, keep_uncorrected=True):
        tris = np.asarray(tris)
        if obs_subset is None:
            obs_subset = np.arange(tris.shape[0])
        if src_subset is None:
            src_subset = np.arange(tris.shape[0])
        self.nearfield = RegularizedNearfieldIntegralOp(
            pts, tris, obs_subset, src_subset, nq_coincident, nq_edge_adj, nq_vert_adjacent,
            nq_far, nq_near, near_threshold, K_near_name, K_far_name, params, float_type,
            plan=near_plan, keep_uncorrected=keep_uncorrected)
        if farfield_op_type is TriToTriDirectFarfieldOp and (farfield_precisions is not None or
                                                             far_plan is not None):
            self.farfield = TriToTriDirectFarfieldOp(nq_far, K_far_name, params, pts, tris,
                                                     float_type, obs_subset, src_subset,
                                                     precisions=farfield_precisions,
                                                     plan=far_plan)
        else:
            self.farfield = farfield_op_type(nq_far, K_far_name, params, pts, tris, float_type,
                                             obs_subset, src_subset)
        self.shape = self.nearfield.shape
        self.device = self.nearfield.device

    def dot_dev(self, v_dev, out=None, prec=None):
        """Device matvec in FP64 storage; ``prec`` selects the farfield
        evaluation ('f32', 'f64' or 'tree') and, when the nearfield keeps an
        FP32 copy, which nearfield blocks are read (FP64 for 'f64')."""
        out = self.farfield.dot_dev(v_dev, out, prec)
        self.nearfield.dot_dev(v_dev, out, accumulate=True, prec=prec or 'f64')
        return out

    def prepare(self, prec):
        """Allocate the lazily created buffers of ``dot_dev(prec)``."""
        if hasattr(self.farfield, 'prepare'):
            self.farfield.prepare(prec)

    def dot(self, v):
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(x).cpu().numpy()

    def nearfield_dot(self, v):
        return self.nearfield.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.nearfield.nearfield_no_correction_dot(v)

    def farfield_dot(self, v):
        return self.farfield.dot(v)
