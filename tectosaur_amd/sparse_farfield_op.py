"""``TriToTriDirectFarfieldOp``: the farfield operator of the summed matvec,
drop-in for tectosaur/ops/sparse_farfield_op.py:16-71.

The reference launches one thread per observation triangle that loops over
every source triangle and every quadrature-point pair
(farfield_tris_to_tris{K}, tectosaur/matrix_free.cl:19-107).  Here the same
sum is computed as three device steps per matvec:

  1. gather   s_j = J_t w_q sum_b phi_b(q) v[9t+3b+:]  (one value per source
               quadrature point; the factor J_t w_q of matrix_free.cl:83 and
               the interpolation of :84-90 folded together),
  2. N-body   r_i = sum_j K(y_j - x_i) s_j             (far_p2p.hip),
  3. scatter  out[9t+3b+:] = J_t sum_q w_q phi_b(q) r_(t,q).

For the point kernels U, T, A, H this is the same sum in a different order
(the kernel of a pair depends only on the two points and normals).  The
surface-curl kernels RT, RA, RH (the vector codes of
tectosaur/kernels/kernels.cl:100-206) depend on whole triangles through the
basis surface curls; they are rewritten as point pairs with richer source
records (the source triangle's curl contracted with the input, folded in by
the gather) and, for RA and RH, 9 moments per observation point that the
scatter contracts with the observation triangle's curls.
"""
import numpy as np
import torch

from . import _native as N
from .farfield import FarfieldEngine, morton_order, prec_of
from .kernels import kernels as KERNELS
from .quadrature import gauss2d_tri


def farfield_plan(nq_far, pts, tris, obs_subset, src_subset, curl):
    """Host-side set-up of the operator (no device work): Morton order of
    the triangles, quadrature points, normals, J w, curls, index maps."""
    from .farfield import group_extent
    pts = np.asarray(pts, dtype=np.float64)
    tris = np.asarray(tris)
    qx, qw = gauss2d_tri(nq_far)
    nq = qx.shape[0]
    phi = np.stack([1 - qx[:, 0] - qx[:, 1], qx[:, 0], qx[:, 1]], axis=1)

    def side(subset):
        v = pts[tris[subset]]
        order = morton_order(v.mean(axis=1))
        v = v[order]
        un = np.cross(v[:, 2] - v[:, 0], v[:, 2] - v[:, 1])
        jac = np.linalg.norm(un, axis=1)
        n = un / jac[:, None]
        qp = np.einsum('qb,tbd->tqd', phi, v).reshape((-1, 3))
        qn = np.repeat(n, nq, axis=0)
        wj = (jac[:, None] * qw[None, :]).reshape(-1)
        return order, qp, qn, wj

    def curls(subset):
        # surface curls of the basis functions, DOF-space order
        # (integral_primitives.cl:95-112): curl_b = (g_b0 e2 - g_b1 e1) / J
        v = pts[tris[subset]]
        e1 = v[:, 1] - v[:, 0]
        e2 = v[:, 2] - v[:, 0]
        J = np.linalg.norm(np.cross(v[:, 2] - v[:, 0], v[:, 2] - v[:, 1]), axis=1)[:, None]
        c = np.stack([(-e2 + e1) / J, e2 / J, -e1 / J], axis=1)
        return np.ascontiguousarray(c.reshape((-1, 9)))

    same = (obs_subset.shape == src_subset.shape and
            np.array_equal(obs_subset, src_subset))
    o_order, o_pts, o_ns, o_wj = side(obs_subset)
    # the full operator (obs set == src set) packs one point set twice
    s_order, s_pts, s_ns, s_wj = (o_order, o_pts, o_ns, o_wj) if same else side(src_subset)
    n_obs = obs_subset.shape[0]
    pt_of_tri = np.empty(n_obs, dtype=np.int32)
    pt_of_tri[o_order] = np.arange(n_obs, dtype=np.int32)
    plan = dict(nq=nq, q=(qx, qw), phi=np.ascontiguousarray(phi), o_pts=o_pts, o_ns=o_ns,
                o_wj=o_wj, s_pts=s_pts, s_ns=s_ns, s_wj=s_wj,
                tri_of_pt=np.repeat(s_order, nq).astype(np.int32), pt_of_tri=pt_of_tri,
                extents=(group_extent(o_pts), group_extent(s_pts)))
    if curl:
        plan['curl_src'] = curls(src_subset)
        plan['curl_obs'] = plan['curl_src'] if same else curls(obs_subset)
    return plan


class TriToTriDirectFarfieldOp:
    def __init__(self, nq_far, K_name, params, pts, tris, float_type, obs_subset, src_subset,
                 tile_cfg=None, precisions=None, plan=None):
        K = KERNELS[K_name]
        self.curl = K.surf_curl_obs or K.surf_curl_src
        obs_subset = np.asarray(obs_subset, dtype=np.int64)
        src_subset = np.asarray(src_subset, dtype=np.int64)
        self.shape = (obs_subset.shape[0] * 9, src_subset.shape[0] * 9)
        self.dim = 3
        self.tensor_dim = K.tensor_dim
        self.n_obs = obs_subset.shape[0]
        self.n_src = src_subset.shape[0]
        self.float_type = float_type
        self.prec = prec_of(float_type)
        self.K_name = K_name
        if plan is None:
            plan = farfield_plan(nq_far, pts, tris, obs_subset, src_subset, self.curl)
        self.q = plan['q']
        self.nq = plan['nq']
        dev = N.require_device()
        self.device = dev
        if precisions is None:
            precisions = (self.prec,)
        self.engine = FarfieldEngine(K_name, params, plan['o_pts'], plan['o_ns'], plan['s_pts'],
                                     plan['s_ns'], precisions=precisions, device=dev,
                                     tile_cfg=tile_cfg, extents=plan['extents'])
        # source point k belongs to DOF-space triangle tri_of_pt[k]
        self.tri_of_pt = N.upload(plan['tri_of_pt'], torch.int32, dev)
        self.pt_of_tri = N.upload(plan['pt_of_tri'], torch.int32, dev)
        self.qphi = N.upload(plan['phi'], torch.float64, dev)
        self.wj_src = N.upload(plan['s_wj'], torch.float64, dev)
        self.wj_obs = N.upload(plan['o_wj'], torch.float64, dev)
        if self.curl:
            self.curl_src = N.upload(plan['curl_src'], torch.float64, dev)
            self.curl_obs = N.upload(plan['curl_obs'], torch.float64, dev)
        self._plan_pts = (plan['o_pts'], plan['s_pts'])
        self.tree = None

    def attach_tree(self, order=4, mac=3.0, pts_per_cell=64, plan=None, n_waves=None):
        """Add a treecode evaluation of this operator (kernels U3, RT3, RA3,
        RH3), used by ``dot_dev(..., prec='tree')`` (FP32) and
        ``tree_dot_dev``; the direct FP32/FP64 paths stay available.
        ``plan``: a precomputed ``fmm.tree_plan`` for these point sets;
        ``n_waves``: target number of list segments (wavefronts)."""
        from .fmm import TreeFarfield
        self.tree = TreeFarfield(self.engine, self._plan_pts[0], self._plan_pts[1], order=order,
                                 mac=mac, pts_per_cell=pts_per_cell, plan=plan, n_waves=n_waves)
        return self.tree

    def tree_dot_dev(self, v_dev, out_dev=None, prec='f32'):
        """Matvec with the treecode in place of the direct N-body."""
        if self.tree is None:
            raise N.TctError('no treecode attached (attach_tree)')
        if out_dev is None:
            out_dev = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        e = self.engine
        self._gather(v_dev, prec)
        part = self.tree.run(prec)
        self._scatter(prec, out_dev, part, self.tree.chunks)
        return out_dev

    def _gather(self, v_dev, prec):
        e = self.engine
        if self.curl:
            e.gather_curl(v_dev, self.tri_of_pt, self.qphi, self.wj_src, self.curl_src, self.nq,
                          prec)
        else:
            e.gather_tris(v_dev, self.tri_of_pt, self.qphi, self.wj_src, self.nq, prec)

    def _scatter(self, prec, out_dev, partial=None, chunks=None):
        e = self.engine
        if e.n_out == 9:
            e.scatter_curl(prec, self.n_obs, self.nq, self.pt_of_tri, self.qphi, self.wj_obs,
                           self.curl_obs, out_dev, partial, chunks)
        else:
            e.scatter_tris(prec, self.n_obs, self.nq, self.pt_of_tri, self.qphi, self.wj_obs,
                           out_dev, partial, chunks)

    @property
    def pair_evals(self):
        return self.engine.pair_evals

    def dot_dev(self, v_dev, out_dev=None, prec=None, nbody_events=None):
        """Device matvec: v_dev float64 [n_src*9] -> float64 [n_obs*9].
        prec: 'f32' or 'f64' (direct N-body), or 'tree' (FP32 treecode when
        one is attached, else the FP32 direct N-body).  nbody_events: an
        optional pair of torch events recorded on the current stream just
        before and after the N-body launch (its duration inside a matvec)."""
        prec = prec or self.prec
        if prec == 'tree':
            if self.tree is not None:
                return self.tree_dot_dev(v_dev, out_dev, 'f32')
            prec = 'f32'
        if out_dev is None:
            out_dev = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        self._gather(v_dev, prec)
        if nbody_events is not None:
            nbody_events[0].record()
        self.engine.run(prec)
        if nbody_events is not None:
            nbody_events[1].record()
        self._scatter(prec, out_dev)
        return out_dev

    def prepare(self, prec):
        """Allocate the buffers a ``dot_dev(prec)`` uses, without running it
        (so a graph capture allocates nothing of its own)."""
        if prec == 'tree' and self.tree is not None:
            self.tree.moments('f32')
            self.tree.partial
        else:
            self.engine.partial('f32' if prec == 'tree' else prec)

    def dot(self, v):
        v_dev = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(v_dev).cpu().numpy()

    async def async_dot(self, v):
        return self.dot(v)

    def nearfield_dot(self, v):
        return self.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.dot(v)

    def farfield_dot(self, v):
        return self.dot(v)
