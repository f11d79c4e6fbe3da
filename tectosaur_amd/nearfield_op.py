"""``RegularizedNearfieldIntegralOp``: the nearfield operator, drop-in for
tectosaur/nearfield/nearfield_op.py:76-168.

Assembly (one-time, per mesh), following nearfield_op.py:95-141:
  * coincident pairs (obs_subset intersect src_subset), integrated with the
    Erichsen-Sauter rule, r2 == 0 points skipped;
  * near pairs |c_i - c_j| <= thr (r_i + r_j) found on the host and split
    by shared vertices into edge-adjacent (rotated so the shared edge is
    local edge 0-1, :44-62), vertex-adjacent (shared vertex rotated to local
    vertex 0) and plain near pairs;
  * each class integrated with K_near and its accurate rule, minus the same
    pairs integrated with K_far and the farfield rule (the farfield operator
    covers every pair with that rule, so the difference corrects it);
  * all blocks placed in one 9x9 BSR matrix on the device (the reference
    keeps four BCOO matrices and applies them on the host).
All integration runs on the device in FP64 (pairs.hip).
"""
import numpy as np
import torch

from . import _native as N
from . import find_near_adj
from .pairs_integrator import PairsIntegrator
from .sparse import bcoo_to_bsr_dev


def edge_adj_orient(tv0, tv1):
    """Rotation that maps the shared edge to local vertices (0, 1)
    (nearfield_op.py:44-50), vectorised: corners {0,1} -> 0, {1,2} -> 1,
    {0,2} -> 2."""
    lo = np.minimum(tv0, tv1)
    hi = np.maximum(tv0, tv1)
    return np.where(lo == 0, np.where(hi == 2, 2, 0), 1)


def resolve_ea_rotation(tris, ea):
    """(obs, src, obs_clicks, src_clicks, src_flip) per edge-adjacent pair
    (nearfield_op.py:52-62)."""
    ea = np.asarray(ea, dtype=np.int64).reshape((-1, 6))
    oc = edge_adj_orient(ea[:, 2], ea[:, 4])
    sc = edge_adj_orient(ea[:, 3], ea[:, 5])
    t = np.asarray(tris)
    flip = ((t[ea[:, 0], (0 + oc) % 3] != t[ea[:, 1], (1 + sc) % 3]) |
            (t[ea[:, 0], (1 + oc) % 3] != t[ea[:, 1], (0 + sc) % 3]))
    return np.stack([ea[:, 0], ea[:, 1], oc, sc, flip.astype(np.int64)], axis=1)


def _is_identity(subset):
    s = np.asarray(subset)
    return s.ndim == 1 and (s.size == 0 or (s[0] == 0 and s[-1] == s.size - 1 and
                                            bool(np.all(np.diff(s) == 1))))


def to_tri_space(dof_indices, obs_subset, src_subset):
    di = np.asarray(dof_indices, dtype=np.int64)
    out = di.copy()
    # identity subsets (the whole mesh) map positions to themselves
    if not _is_identity(obs_subset):
        out[:, 0] = obs_subset[di[:, 0]]
    if not _is_identity(src_subset):
        out[:, 1] = src_subset[di[:, 1]]
    return out


def to_dof_space(tri_indices, obs_subset, src_subset):
    """Positions of (obs tri, src tri) within the subsets (nearfield_op.py:24-38)."""
    ti = np.asarray(tri_indices, dtype=np.int64).reshape((-1, 2))
    n = int(max(obs_subset.max(initial=-1), src_subset.max(initial=-1))) + 1
    opos = np.full(n, -1, dtype=np.int64)
    spos = np.full(n, -1, dtype=np.int64)
    opos[obs_subset] = np.arange(obs_subset.shape[0])
    spos[src_subset] = np.arange(src_subset.shape[0])
    out = np.stack([opos[ti[:, 0]], spos[ti[:, 1]]], axis=1) if ti.shape[0] else \
        np.zeros((0, 2), dtype=np.int64)
    if (out < 0).any():
        raise ValueError('pair outside the obs/src subsets')
    return out


def nearfield_plan(pts, tris, obs_subset, src_subset, near_threshold, finder='host'):
    """Pair classification: coincident, edge-adjacent (rotated),
    vertex-adjacent and near pairs, in triangle and DOF space
    (nearfield_op.py:95-114).  finder 'host' searches with the C++ finder
    (no device work, safe on worker threads); 'device' with the device
    finder and classifier (the same pairs, in the same order)."""
    pts = np.asarray(pts, dtype=np.float64)
    tris = np.asarray(tris, dtype=np.int64)
    obs_subset = np.asarray(obs_subset, dtype=np.int64)
    src_subset = np.asarray(src_subset, dtype=np.int64)
    co_tris = np.intersect1d(obs_subset, src_subset)
    co_indices = np.stack([co_tris, co_tris], axis=1)
    co_dofs = to_dof_space(co_indices, obs_subset, src_subset)
    if finder == 'device':
        import torch
        ta, tb = tris[obs_subset], tris[src_subset]
        close = find_near_adj.find_near_pairs_dev(
            *find_near_adj.get_tri_centroids_rs(pts, ta),
            *find_near_adj.get_tri_centroids_rs(pts, tb), near_threshold)
        d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=close.device)
        near_dofs, va_dofs, ea_dofs = (x.cpu().numpy() for x in
                                       find_near_adj.split_adjacent_close_dev(close, d(ta), d(tb)))
    elif finder == 'host':
        close = find_near_adj.find_close_or_touching(pts, tris[obs_subset], pts,
                                                     tris[src_subset], near_threshold)
        near_dofs, va_dofs, ea_dofs = find_near_adj.split_adjacent_close(
            close, tris[obs_subset], tris[src_subset])
    else:
        raise ValueError("finder must be 'host' or 'device'")
    near = to_tri_space(near_dofs, obs_subset, src_subset)
    va = to_tri_space(va_dofs, obs_subset, src_subset)
    va = np.concatenate([va, np.zeros((va.shape[0], 1), dtype=np.int64)], axis=1)
    ea = resolve_ea_rotation(tris, to_tri_space(ea_dofs, obs_subset, src_subset))
    return dict(co=(co_indices, co_dofs), ea=(ea, ea_dofs[:, :2]), va=(va, va_dofs[:, :2]),
                near=(near, near_dofs))


class RegularizedNearfieldIntegralOp:
    def __init__(self, pts, tris, obs_subset, src_subset, nq_coincident, nq_edge_adj,
                 nq_vert_adjacent, nq_far, nq_near, near_threshold, K_near_name, K_far_name,
                 params, float_type, keep_classes=False, plan=None, keep_uncorrected=True):
        pts = np.asarray(pts, dtype=np.float64)
        tris = np.asarray(tris, dtype=np.int64)
        obs_subset = np.asarray(obs_subset, dtype=np.int64)
        src_subset = np.asarray(src_subset, dtype=np.int64)
        self.shape = (obs_subset.shape[0] * 9, src_subset.shape[0] * 9)
        self.device = N.require_device()
        dev = self.device
        pairs_int = PairsIntegrator(K_near_name, params, float_type, nq_far, nq_near, pts, tris)
        corr_int = PairsIntegrator(K_far_name, params, float_type, nq_far, nq_near, pts, tris)
        if plan is None:
            plan = nearfield_plan(pts, tris, obs_subset, src_subset, near_threshold)

        from . import triangle_rules
        from .quadrature import gauss4d_tri
        nqv = nq_vert_adjacent if not isinstance(nq_vert_adjacent, (int, np.integer)) else \
            (nq_vert_adjacent,) * 3
        rules = {
            'co': pairs_int._rule(('co', nq_coincident),
                                  lambda: triangle_rules.coincident_quad(nq_coincident)),
            'ea': pairs_int._rule(('ea', nq_edge_adj),
                                  lambda: triangle_rules.edge_adj_quad(nq_edge_adj)),
            'va': pairs_int._rule(('va',) + tuple(nqv),
                                  lambda: triangle_rules.vertex_adj_quad(*nqv)),
            'near': pairs_int.near_q,
        }
        far_q = corr_int.far_q
        classes = (('co',) + plan['co'] + (True, True),
                   ('ea',) + plan['ea'] + (True, False),
                   ('va',) + plan['va'] + (True, False),
                   ('near',) + plan['near'] + (False, False))
        rows, cols, corrected, uncorrected = [], [], [], []
        self.counts = {}
        self.classes = {}
        for name, prs, dofs, check0, corr_check0 in classes:
            self.counts[name] = int(prs.shape[0])
            if prs.shape[0] == 0:
                continue
            from .pairs_integrator import check_pairs
            check_pairs(prs, tris.shape[0])
            p_dev = N.upload(prs.astype(np.int32), None, dev)
            acc = pairs_int.pairs_quad_dev(rules[name], p_dev, check0)
            corr = corr_int.pairs_quad_dev(far_q, p_dev[:, :2].contiguous(), corr_check0)
            diff = torch.empty_like(acc)
            N.call('tct_sub_f64', acc.numel(), N.ptr(acc), N.ptr(corr), N.ptr(diff),
                   N.stream_ptr())
            rows.append(N.upload(dofs[:, 0], None, dev))
            cols.append(N.upload(dofs[:, 1], None, dev))
            corrected.append(diff)
            if keep_uncorrected:
                uncorrected.append(acc)
            if keep_classes:
                self.classes[name] = (prs, dofs, acc, corr)
            del corr
        if rows:
            r = torch.cat(rows)
            c = torch.cat(cols)
            self.mat = bcoo_to_bsr_dev(r, c, torch.cat(corrected), self.shape)
            self.mat_no_correction = (bcoo_to_bsr_dev(r, c, torch.cat(uncorrected), self.shape)
                                      if keep_uncorrected else None)
        else:
            z = torch.zeros((0, 81), dtype=torch.float64, device=dev)
            e = torch.zeros(0, dtype=torch.int64, device=dev)
            self.mat = bcoo_to_bsr_dev(e, e, z, self.shape)
            self.mat_no_correction = self.mat
        self.n_pairs = sum(self.counts.values())

    def dot_dev(self, x, out=None, accumulate=False, prec='f64'):
        """prec other than 'f64' uses the FP32 copy of the blocks when one
        exists (``use_f32_inner``)."""
        return self.mat.dot_dev(x, out, accumulate, prec)

    def use_f32_inner(self):
        """Keep an FP32 copy of the corrected blocks for inexact (inner)
        matvecs of mixed-precision solves; FP64 products are unchanged."""
        self.mat.make_f32()
        return self

    def dot(self, v):
        return self.mat.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.mat_no_correction.dot(v)

    def to_dense(self):
        return self.mat.to_dense()

    def no_correction_to_dense(self):
        return self.mat_no_correction.to_dense()

    def full_scipy_mat(self):
        return self.mat.to_scipy()

    def full_scipy_mat_no_correction(self):
        return self.mat_no_correction.to_scipy()
