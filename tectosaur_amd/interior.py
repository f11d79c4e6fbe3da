"""Interior evaluation: the boundary operators at observation points off (or
on) the mesh, drop-in for tectosaur/interior.py.

``InteriorOp(obs_pts, obs_ns, src_mesh, K_name, threshold, nq_vertex, nq_far,
nq_near, params, float_type)`` keeps the reference's decomposition
(tectosaur/interior.py:29-119):

    dot(v) = far(v)                  every (point, triangle) pair, nq_far rule
           + near(v)                 close pairs, nq_near rule
           - near_correction(v)      every close pair, nq_far rule
           + vertex(v)               points at a triangle vertex, corner rule

with "close" = |x - c_t| <= 4 r_t (the reference fixes the threshold at 4.0
whatever the argument, :33) and vertex pairs split off by a 1e-9 distance
test (fast_find_nearfield.cpp:184-215).

On the device:
  * far(v) is the farfield N-body of far_p2p.hip with the observation points
    as observers: the gather folds J w phi into per-quadrature-point source
    strengths (the surface curls for RT), the kernel sums, and the reduction
    writes one 3-vector per point -- farfield_tris_to_pts{K}
    (tectosaur/matrix_free.cl:109-194) in another order.
  * the three pair terms are one 3 x 9 block per close pair, formed once at
    set-up from the FP64 pair integrals of interior.hip (near or corner
    block minus the far-rule block) and applied with tct_bsr39_spmv.
"""
import numpy as np
import scipy.sparse
import torch

from . import _native as N
from .farfield import FarfieldEngine, morton_order, prec_of, group_extent
from .find_near_adj import find_near_pairs, get_tri_centroids_rs
from .kernels import kernels as KERNELS
from .quadrature import gauss2d_tri, gaussxw, map_to

# the reference's InteriorOp ignores its threshold argument and uses this
# (tectosaur/interior.py:33)
INTERIOR_THRESHOLD = 4.0
VERTEX_TOL = 1e-9


def _pt_kernel_ok(K):
    if K.surf_curl_obs:
        raise N.TctError('%s has no point-observer form (observation surface curl)' % K.name)


def split_vertex_nearfield(pairs, obs_pts, src_pts, src_tris):
    """(vertex pairs [n,3] = (obs, tri, vertex), near pairs [n,3] = (obs,
    tri, 0)), each in input order; the vertex is the last one of the
    triangle within 1e-9 of the point (fast_find_nearfield.cpp:184-215)."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape((-1, 2))
    if pairs.shape[0] == 0:
        z = np.zeros((0, 3), dtype=np.int64)
        return z, z.copy()
    x = np.asarray(obs_pts, dtype=np.float64)[pairs[:, 0]]
    v = np.asarray(src_pts, dtype=np.float64)[np.asarray(src_tris)[pairs[:, 1]]]
    hit = np.linalg.norm(v - x[:, None, :], axis=2) < VERTEX_TOL
    isv = np.where(hit[:, 2], 2, np.where(hit[:, 1], 1, np.where(hit[:, 0], 0, -1)))
    out = np.concatenate([pairs, np.maximum(isv, 0)[:, None]], axis=1)
    return out[isv >= 0], out[isv < 0]


def make_pairs_mat(pairs, entries, shape):
    """Host CSR of 3 x 9 pair blocks, tectosaur/interior.py:14-26."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape((-1, 3))
    e = np.asarray(entries, dtype=np.float64).reshape((-1, 3, 9))
    rows = np.broadcast_to((pairs[:, 0] * 3)[:, None, None] + np.arange(3)[None, :, None],
                           e.shape)
    cols = np.broadcast_to((pairs[:, 1] * 9)[:, None, None] + np.arange(9)[None, None, :],
                           e.shape)
    return scipy.sparse.csr_matrix((e.reshape(-1), (rows.reshape(-1), cols.reshape(-1))),
                                   shape=shape)


def _dev_pairs_quad(kid, pairs_dev, n_pairs, quad, obs_pts, obs_ns, src_pts, src_tris, params,
                    finite_part=False, check0=False):
    out = torch.empty((n_pairs, 27), dtype=torch.float64, device=obs_pts.device)
    N.call('tct_interior_pairs', kid, quad[1].shape[0], N.ptr(quad[0]), N.ptr(quad[1]),
           N.ptr(obs_pts), N.ptr(obs_ns), N.ptr(src_pts), N.ptr(src_tris), n_pairs,
           N.ptr(pairs_dev) if n_pairs else None, int(check0), int(bool(finite_part)),
           params.ctypes.data, 0, N.ptr(out), N.stream_ptr())
    return out


def _dev_pairs_sinh(kid, pairs_dev, n_pairs, rule01, obs_pts, obs_ns, src_pts, src_tris,
                    params):
    out = torch.empty((n_pairs, 27), dtype=torch.float64, device=obs_pts.device)
    n = rule01[1].shape[0]
    N.call('tct_interior_pairs_sinh', kid, n, N.ptr(rule01[0]), N.ptr(rule01[1]), n,
           N.ptr(rule01[0]), N.ptr(rule01[1]), N.ptr(obs_pts), N.ptr(obs_ns), N.ptr(src_pts),
           N.ptr(src_tris), n_pairs, N.ptr(pairs_dev) if n_pairs else None, params.ctypes.data,
           N.ptr(out), N.stream_ptr())
    return out


def _dev_corners(pairs_dev, n_pairs, quad, obs_ns, src_pts, src_tris, params):
    out = torch.empty((n_pairs, 27), dtype=torch.float64, device=obs_ns.device)
    N.call('tct_interior_corners', quad[1].shape[0], N.ptr(quad[0]), N.ptr(quad[1]),
           N.ptr(obs_ns), N.ptr(src_pts), N.ptr(src_tris), n_pairs,
           N.ptr(pairs_dev) if n_pairs else None, params.ctypes.data, N.ptr(out),
           N.stream_ptr())
    return out


def _quad_dev(q, dev):
    return tuple(torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
                 for a in q)


def interior_pairs_quad(K_name, pairs_list, quad, obs_pts, obs_ns, src_pts, src_tris, params,
                        float_type=np.float64, finite_part=False):
    """Pair integrals [n, 3, 3, 3] (tectosaur/interior.py:121-142); host
    arrays in and out, computed in FP64 whatever float_type."""
    K = KERNELS[K_name]
    _pt_kernel_ok(K)
    dev = N.require_device()
    f64 = dict(dtype=torch.float64, device=dev)
    pairs = np.ascontiguousarray(np.asarray(pairs_list).reshape((-1, 3)), dtype=np.int32)
    n = pairs.shape[0]
    if n == 0:
        return np.zeros((0, 3, 3, 3), dtype=float_type)
    out = _dev_pairs_quad(
        K.kid, torch.as_tensor(pairs, device=dev), n, _quad_dev(quad, dev),
        torch.as_tensor(np.ascontiguousarray(obs_pts, dtype=np.float64), **f64),
        torch.as_tensor(np.ascontiguousarray(obs_ns, dtype=np.float64), **f64),
        torch.as_tensor(np.ascontiguousarray(src_pts, dtype=np.float64), **f64),
        torch.as_tensor(np.ascontiguousarray(src_tris, dtype=np.int32), device=dev),
        N.params_array(params), finite_part)
    return out.cpu().numpy().reshape((n, 3, 3, 3)).astype(float_type)


class TriToPtDirectFarfieldOp:
    """out[3i + d] = sum over every source triangle and quadrature point of
    K(y_q - x_i) (J w phi . v), pairs at r2 == 0 skipped: the reference's
    TriToPtDirectFarfieldOp (tectosaur/interior.py:139-185) as gather +
    N-body + point reduction.  ``dot_dev`` writes a device tensor
    [n_obs * 3] in the caller's point order."""

    def __init__(self, obs_pts, obs_ns, src_mesh, K_name, nq, params, float_type,
                 precisions=None, device=None):
        K = KERNELS[K_name]
        _pt_kernel_ok(K)
        self.K = K
        src_pts = np.asarray(src_mesh[0], dtype=np.float64)
        src_tris = np.asarray(src_mesh[1], dtype=np.int64)
        obs_pts = np.asarray(obs_pts, dtype=np.float64).reshape((-1, 3))
        obs_ns = np.asarray(obs_ns, dtype=np.float64).reshape((-1, 3))
        self.shape = (obs_pts.shape[0] * 3, src_tris.shape[0] * 9)
        self.dim = 3
        self.tensor_dim = K.tensor_dim
        self.n_obs = obs_pts.shape[0]
        self.n_src = src_tris.shape[0]
        self.float_type = float_type
        self.prec = prec_of(float_type)
        self.q = gauss2d_tri(nq)
        qx, qw = self.q
        self.nq = qw.shape[0]
        phi = np.stack([1 - qx[:, 0] - qx[:, 1], qx[:, 0], qx[:, 1]], axis=1)
        v = src_pts[src_tris]
        s_order = morton_order(v.mean(axis=1))
        vs = v[s_order]
        un = np.cross(vs[:, 2] - vs[:, 0], vs[:, 2] - vs[:, 1])
        jac = np.linalg.norm(un, axis=1)
        s_pts = np.einsum('qb,tbd->tqd', phi, vs).reshape((-1, 3))
        s_ns = np.repeat(un / jac[:, None], self.nq, axis=0)
        s_wj = (jac[:, None] * qw[None, :]).reshape(-1)
        self.po = morton_order(obs_pts)
        o_pts = obs_pts[self.po]
        self.device = device if device is not None else N.require_device()
        dev = self.device
        self.engine = FarfieldEngine(K_name, params, o_pts, obs_ns[self.po], s_pts, s_ns,
                                     precisions=precisions or (self.prec,), device=dev,
                                     extents=(group_extent(o_pts), group_extent(s_pts)))
        i32 = dict(dtype=torch.int32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.tri_of_pt = torch.as_tensor(np.repeat(s_order, self.nq).astype(np.int32), **i32)
        self.qphi = torch.as_tensor(np.ascontiguousarray(phi), **f64)
        self.wj_src = torch.as_tensor(s_wj, **f64)
        self.curl_src = None
        if K.surf_curl_src:
            e1 = v[:, 1] - v[:, 0]
            e2 = v[:, 2] - v[:, 0]
            J = np.linalg.norm(np.cross(v[:, 2] - v[:, 0], v[:, 2] - v[:, 1]), axis=1)[:, None]
            c = np.stack([(-e2 + e1) / J, e2 / J, -e1 / J], axis=1).reshape((-1, 9))
            self.curl_src = torch.as_tensor(np.ascontiguousarray(c), **f64)
        # point k of the packed observation set is caller point po[k]
        self.po_dev = torch.as_tensor(self.po, dtype=torch.int64, device=dev)
        self._packed = torch.empty((self.n_obs, 3), **f64)

    @property
    def pair_evals(self):
        return self.engine.pair_evals

    def dot_dev(self, v_dev, out=None, prec=None):
        prec = prec or self.prec
        e = self.engine
        if self.curl_src is not None:
            e.gather_curl(v_dev, self.tri_of_pt, self.qphi, self.wj_src, self.curl_src, self.nq,
                          prec)
        else:
            e.gather_tris(v_dev, self.tri_of_pt, self.qphi, self.wj_src, self.nq, prec)
        e.run(prec)
        e.reduce_pts(prec, self._packed)
        if out is None:
            out = torch.empty(self.n_obs * 3, dtype=torch.float64, device=self.device)
        out.view(-1, 3).index_copy_(0, self.po_dev, self._packed)
        return out

    def dot(self, v):
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(x).cpu().numpy()


class InteriorOp:
    """The boundary operator K_name of a source mesh evaluated at
    observation points (tectosaur/interior.py:29-119); same signature and
    decomposition.  ``float_type`` selects the farfield N-body precision;
    the pair blocks are FP64.  ``dot`` takes and returns host arrays,
    ``dot_dev`` device tensors.

    near_method 'gauss' integrates close pairs with the nq_near triangle
    rule, as the reference does; 'sinh' with the polar sinh rule of
    sinh_order x sinh_order points per sub-triangle, accurate for points
    arbitrarily close to (but off) the surface (no reference
    counterpart)."""

    def __init__(self, obs_pts, obs_ns, src_mesh, K_name, threshold, nq_vertex, nq_far, nq_near,
                 params, float_type, near_method='gauss', sinh_order=24):
        K = KERNELS[K_name]
        _pt_kernel_ok(K)
        self.K_name = K_name
        self.float_type = float_type
        self.threshold = INTERIOR_THRESHOLD
        obs_pts = np.ascontiguousarray(obs_pts, dtype=np.float64).reshape((-1, 3))
        obs_ns = np.ascontiguousarray(obs_ns, dtype=np.float64).reshape((-1, 3))
        src_pts = np.ascontiguousarray(src_mesh[0], dtype=np.float64)
        src_tris = np.ascontiguousarray(src_mesh[1], dtype=np.int64)
        n_obs = obs_pts.shape[0]
        close = find_near_pairs(obs_pts, np.zeros(n_obs), *get_tri_centroids_rs(src_pts, src_tris),
                                self.threshold)
        self.vertex_pairs, self.near_pairs = split_vertex_nearfield(close, obs_pts, src_pts,
                                                                    src_tris)
        self.all_near_pairs = np.vstack((self.near_pairs, self.vertex_pairs))
        self.all_near_pairs[:, 2] = 0

        self.farfield = TriToPtDirectFarfieldOp(obs_pts, obs_ns, (src_pts, src_tris), K_name,
                                                nq_far, params, float_type)
        self.shape = self.farfield.shape
        self.device = dev = self.farfield.device
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        params_a = N.params_array(params)
        g_obs = torch.as_tensor(obs_pts, **f64)
        g_ns = torch.as_tensor(obs_ns, **f64)
        g_pts = torch.as_tensor(src_pts, **f64)
        g_tris = torch.as_tensor(src_tris.astype(np.int32), **i32)
        far_q = _quad_dev(gauss2d_tri(nq_far), dev)
        near_q = _quad_dev(gauss2d_tri(nq_near), dev)
        vert_q = _quad_dev(map_to(gaussxw(nq_vertex), [0.0, np.pi / 2]), dev)
        if near_method not in ('gauss', 'sinh'):
            raise ValueError("near_method must be 'gauss' or 'sinh'")
        sinh_q = _quad_dev(map_to(gaussxw(sinh_order), [0.0, 1.0]), dev)
        self.near_method = near_method

        # one block per close pair, in (point, triangle) order: the near-rule
        # or corner block minus the far-rule block the farfield also added
        n = close.shape[0]
        self.n_blocks = n
        is_vert = np.zeros(n, dtype=bool)
        clicks = np.zeros(n, dtype=np.int32)
        if n:
            key = close[:, 0] * src_tris.shape[0] + close[:, 1]
            vkey = self.vertex_pairs[:, 0] * src_tris.shape[0] + self.vertex_pairs[:, 1]
            vpos = np.searchsorted(key, vkey)
            is_vert[vpos] = True
            clicks[vpos] = self.vertex_pairs[:, 2]
        blocks = torch.empty((n, 27), **f64)
        if n:
            trip = np.ascontiguousarray(np.stack([close[:, 0], close[:, 1], clicks], axis=1),
                                        dtype=np.int32)
            trip_d = torch.as_tensor(trip, device=dev)
            ni = np.nonzero(~is_vert)[0]
            vi = np.nonzero(is_vert)[0]
            if ni.shape[0]:
                sel = torch.as_tensor(ni, dtype=torch.int64, device=dev)
                sub = trip_d.index_select(0, sel).contiguous()
                if near_method == 'sinh':
                    near_b = _dev_pairs_sinh(K.kid, sub, ni.shape[0], sinh_q, g_obs, g_ns, g_pts,
                                             g_tris, params_a)
                else:
                    near_b = _dev_pairs_quad(K.kid, sub, ni.shape[0], near_q, g_obs, g_ns, g_pts,
                                             g_tris, params_a)
                blocks.index_copy_(0, sel, near_b)
            if vi.shape[0]:
                sel = torch.as_tensor(vi, dtype=torch.int64, device=dev)
                blocks.index_copy_(0, sel, _dev_corners(
                    trip_d.index_select(0, sel).contiguous(), vi.shape[0], vert_q, g_ns, g_pts,
                    g_tris, params_a))
            trip0 = trip.copy()
            trip0[:, 2] = 0
            blocks.sub_(_dev_pairs_quad(K.kid, torch.as_tensor(trip0, device=dev), n, far_q,
                                        g_obs, g_ns, g_pts, g_tris, params_a))
        indptr = np.zeros(n_obs + 1, dtype=np.int32)
        if n:
            np.cumsum(np.bincount(close[:, 0], minlength=n_obs), out=indptr[1:])
        self.indptr = torch.as_tensor(indptr, **i32)
        self.indices = torch.as_tensor(close[:, 1].astype(np.int32), **i32)
        self.blocks = blocks

    @property
    def pair_evals(self):
        return self.farfield.pair_evals

    def dot_dev(self, v_dev, out=None, prec=None):
        out = self.farfield.dot_dev(v_dev, out, prec)
        N.call('tct_bsr39_spmv', self.farfield.n_obs, N.ptr(self.indptr), N.ptr(self.indices),
               N.ptr(self.blocks), N.ptr(v_dev), N.ptr(out), 1, N.stream_ptr())
        return out

    def dot(self, v):
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(x).cpu().numpy()

    def pair_blocks(self):
        """(pairs [n,2] (point, triangle), corrected blocks [n,3,9]) on the
        host."""
        pairs = np.stack([np.repeat(np.arange(self.farfield.n_obs),
                                    np.diff(self.indptr.cpu().numpy())),
                          self.indices.cpu().numpy()], axis=1)
        return pairs, self.blocks.cpu().numpy().reshape((-1, 3, 9))
