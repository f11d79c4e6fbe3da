"""Farfield direct N-body on the device.

``FarfieldEngine`` owns the packed device layout of one (observation
points, source points) pair of point sets and launches the HIP kernels of
far_p2p.hip.  ``farfield_pts_direct`` mirrors the reference entry point of
the same name (tectosaur/farfield.py:15-38).

Point sets are packed in Morton (Z-curve) order so each group of 64
consecutive points is spatially compact: the FP32 path stores coordinates
relative to an FP64 origin per group, and its accuracy depends on the group
extent.
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from .kernels import kernels as KERNELS

GROUP = N.GROUP
REC = N.REC
U32 = 2.0 ** -24  # FP32 unit roundoff


def morton_order(p):
    """Permutation sorting points [n, 3] along a 3-D Z-curve (21 bits/axis)."""
    p = np.asarray(p, dtype=np.float64)
    if p.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    lo = p.min(axis=0)
    ext = np.max(p.max(axis=0) - lo)
    ext = ext if ext > 0 else 1.0
    q = np.clip(((p - lo) / ext * (2 ** 21 - 1)).astype(np.int64), 0, 2 ** 21 - 1)

    def spread(v):
        v = v & 0x1FFFFF
        v = (v | (v << 32)) & 0x1F00000000FFFF
        v = (v | (v << 16)) & 0x1F0000FF0000FF
        v = (v | (v << 8)) & 0x100F00F00F00F00F
        v = (v | (v << 4)) & 0x10C30C30C30C30C3
        v = (v | (v << 2)) & 0x1249249249249249
        return v

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    return np.argsort(code, kind='stable')


def round_up(n, m):
    return ((n + m - 1) // m) * m


def group_extent(p):
    """Largest distance of a point from the mean of its 64-point group."""
    p = np.asarray(p, dtype=np.float64)
    n = p.shape[0]
    if n == 0:
        return 0.0
    q = np.full((round_up(n, GROUP), 3), np.nan)
    q[:n] = p
    q = q.reshape((-1, GROUP, 3))
    cnt = np.sum(~np.isnan(q[:, :, 0]), axis=1)
    mean = np.nansum(q, axis=1) / cnt[:, None]
    d = np.linalg.norm(q - mean[:, None, :], axis=2)
    return float(np.nanmax(d))


def default_tile_cfg(n_obs_pad, n_src_pad, prec, device=None):
    """Launch shape when no tuned entry exists: about 32 workgroups per CU,
    the source list split into chunks until the observation blocks reach
    that count (the shape the tuner picks across the served size range:
    block 128 below 8k observation points, else 256, one point per lane)."""
    # a fixed CU count (MI355X: 256), not the device's: the chunk count
    # sets the summation order, so results do not depend on the box
    n_cu = 256
    block = 256 if n_obs_pad >= 8192 else 128
    per_lane = 1
    obs_blocks = -(-n_obs_pad // (block * per_lane))
    n_tiles = n_src_pad // GROUP
    chunks = max(1, min(n_tiles, -(-32 * n_cu // obs_blocks)))
    return N.TileCfg(block, per_lane, chunks, 0)


class FarfieldEngine:
    """Device layout + launches for one obs/src point-set pair and kernel.

    obs_pts, obs_ns, src_pts, src_ns: float64 [n, 3] numpy arrays, already
    in packing order.  ``precisions`` lists the layouts to build ('f32',
    'f64').
    """

    def __init__(self, K_name, params, obs_pts, obs_ns, src_pts, src_ns,
                 precisions=('f32', 'f64'), device=None, tile_cfg=None, extents=None):
        K = KERNELS[K_name]
        self.K = K
        # record words per source point and sums per observation point
        self.rec_words = REC
        self.n_out = 9 if K_name in ('elasticRA3', 'elasticRH3') else 3
        self.kid = K.kid
        self.params = N.params_array(params)
        self.device = device if device is not None else N.require_device()
        self.n_obs = obs_pts.shape[0]
        self.n_src = src_pts.shape[0]
        if self.n_obs == 0 or self.n_src == 0:
            raise N.TctError('empty point set')
        self.n_obs_pad = round_up(self.n_obs, GROUP)
        self.n_src_pad = round_up(self.n_src, GROUP)
        self.precisions = tuple(precisions)
        dev = self.device
        f64 = dict(dtype=torch.float64, device=dev)
        st = N.stream_ptr()
        op = N.upload(obs_pts, torch.float64, dev)
        on = N.upload(obs_ns, torch.float64, dev)
        sp = N.upload(src_pts, torch.float64, dev)
        sn = N.upload(src_ns, torch.float64, dev)
        self.loc32 = torch.empty(6 * self.n_obs_pad, dtype=torch.float32, device=dev)
        self.obs_org = torch.empty(3 * (self.n_obs_pad // GROUP), **f64)
        self.obs_soa64 = torch.empty(6 * self.n_obs_pad, **f64)
        N.call('tct_far_pack_obs', self.n_obs, N.ptr(op), N.ptr(on), self.n_obs_pad,
               N.ptr(self.loc32), N.ptr(self.obs_org), N.ptr(self.obs_soa64), st)
        self.rec32 = torch.empty(self.rec_words * self.n_src_pad, dtype=torch.float32, device=dev)
        self.src_org = torch.empty(3 * (self.n_src_pad // GROUP), **f64)
        self.rec64 = torch.empty(self.rec_words * self.n_src_pad, **f64)
        N.call('tct_far_pack_src_ex', self.n_src, N.ptr(sp), N.ptr(sn), self.n_src_pad,
               self.rec_words, N.ptr(self.rec32), N.ptr(self.src_org), N.ptr(self.rec64), st)
        # rounding bound of one point stored against two different origins
        if extents is None:
            extents = (group_extent(obs_pts), group_extent(src_pts))
        self.group_extent = float(extents[0] + extents[1])
        self.eps2 = (8.0 * U32 * self.group_extent) ** 2
        self.tile_cfg = {}
        for prec in ('f32', 'f64'):
            cfg = tile_cfg.get(prec) if isinstance(tile_cfg, dict) else tile_cfg
            if cfg is None:
                from . import autotune
                cfg = autotune.lookup(self.K.name, prec, self.n_obs_pad, self.n_src_pad,
                                      self.device)
            self.tile_cfg[prec] = cfg
        self._partial = {}

    def _prec_id(self, prec):
        if prec not in ('f32', 'f64'):
            raise ValueError('precision must be f32 or f64')
        return N.F32 if prec == 'f32' else N.F64

    def rec(self, prec):
        return self.rec32 if prec == 'f32' else self.rec64

    def partial(self, prec):
        cfg = self.tile_cfg[prec]
        key = (prec, cfg.src_chunks)
        if key not in self._partial:
            self._partial[key] = torch.empty(cfg.src_chunks * self.n_out * self.n_obs_pad,
                                             dtype=torch.float64, device=self.device)
        return self._partial[key]

    def set_strengths(self, vec_dev, prec):
        """vec_dev: float64 device tensor [n_src, 3] in packing order."""
        if self.K.surf_curl_obs or self.K.surf_curl_src:
            raise N.TctError('%s has no point-to-point form' % self.K.name)
        N.call('tct_far_strength_pts', self.kid, self._prec_id(prec), self.n_src,
               N.ptr(vec_dev), self.params.ctypes.data, N.ptr(self.rec(prec)), N.stream_ptr())

    def gather_tris(self, v_dev, tri_of_pt, qphi, wj, nq, prec):
        N.call('tct_far_gather_tris', self.kid, self._prec_id(prec), self.n_src, nq,
               N.ptr(tri_of_pt), N.ptr(qphi), N.ptr(wj), N.ptr(v_dev), self.params.ctypes.data,
               N.ptr(self.rec(prec)), N.stream_ptr())

    def gather_curl(self, v_dev, tri_of_pt, qphi, wj, tri_curl, nq, prec):
        N.call('tct_far_gather_curl', self.kid, self._prec_id(prec), self.n_src, nq,
               N.ptr(tri_of_pt), N.ptr(qphi), N.ptr(wj), N.ptr(tri_curl), N.ptr(v_dev),
               self.params.ctypes.data, N.ptr(self.rec(prec)), N.stream_ptr())

    def scatter_curl(self, prec, n_tris, nq, pt_of_tri, qphi, wj_obs, tri_curl, out_dev,
                     partial=None, chunks=None):
        """``partial``/``chunks`` override the N-body's own buffer (the
        treecode writes a buffer of its own with its own chunk count)."""
        chunks = chunks or self.tile_cfg[prec].src_chunks
        partial = partial if partial is not None else self.partial(prec)
        N.call('tct_far_scatter_curl', self.kid, n_tris, nq, N.ptr(pt_of_tri), N.ptr(qphi),
               N.ptr(wj_obs), N.ptr(tri_curl), self.n_obs_pad, chunks, N.ptr(partial),
               self.params.ctypes.data, N.ptr(out_dev), N.stream_ptr())

    def run(self, prec, cfg=None):
        """Launch the N-body kernel; returns the partial-sum buffer."""
        if cfg is not None:
            self.tile_cfg[prec] = cfg
        cfg = self.tile_cfg[prec]
        out = self.partial(prec)
        N.call('tct_far_p2p', self.kid, self._prec_id(prec), self.n_obs_pad, N.ptr(self.loc32),
               N.ptr(self.obs_org), N.ptr(self.obs_soa64), self.n_src_pad,
               N.ptr(self.rec(prec)), N.ptr(self.src_org), self.params.ctypes.data,
               float(self.eps2), self.group_extent, ctypes_ref(cfg), N.ptr(out),
               N.stream_ptr())
        return out

    def reduce_pts(self, prec, out_dev, partial=None, chunks=None):
        chunks = chunks or self.tile_cfg[prec].src_chunks
        partial = partial if partial is not None else self.partial(prec)
        N.call('tct_far_reduce_pts', self.n_obs, self.n_obs_pad, chunks, N.ptr(partial),
               N.ptr(out_dev), N.stream_ptr())

    def scatter_tris(self, prec, n_tris, nq, pt_of_tri, qphi, wj_obs, out_dev, partial=None,
                     chunks=None):
        chunks = chunks or self.tile_cfg[prec].src_chunks
        partial = partial if partial is not None else self.partial(prec)
        N.call('tct_far_scatter_tris', n_tris, nq, N.ptr(pt_of_tri), N.ptr(qphi), N.ptr(wj_obs),
               self.n_obs_pad, chunks, N.ptr(partial), N.ptr(out_dev), N.stream_ptr())

    @property
    def pair_evals(self):
        return self.n_obs * self.n_src


def run_batched(engines, block=256, obs_per_lane=1):
    """One launch for the FP32 N-bodies of several engines (same kernel and
    constants), each with its own layout and source chunks; every engine's
    partial-sum buffer gets what its own ``run('f32')`` would write.
    Returns the partial-sum buffers."""
    if not engines:
        return []
    e0 = engines[0]
    for e in engines[1:]:
        if e.kid != e0.kid or not np.array_equal(e.params, e0.params):
            raise N.TctError('batched N-body: engines differ in kernel or constants')
    segs = (N.FarSegment * len(engines))()
    outs = []
    for i, e in enumerate(engines):
        cfg = e.tile_cfg['f32']
        out = e.partial('f32')
        outs.append(out)
        segs[i] = N.FarSegment(e.n_obs_pad, N.ptr(e.loc32), N.ptr(e.obs_org), e.n_src_pad,
                               N.ptr(e.rec32), N.ptr(e.src_org), float(e.eps2),
                               e.group_extent, cfg.src_chunks, cfg.flags & 1, N.ptr(out))
    N.call('tct_far_p2p_batched', e0.kid, e0.params.ctypes.data, block, obs_per_lane,
           len(engines), ctypes.addressof(segs), N.stream_ptr())
    return outs


def ctypes_ref(cfg):
    import ctypes
    return ctypes.byref(cfg)


def prec_of(float_type):
    return 'f32' if np.dtype(float_type) == np.float32 else 'f64'


def farfield_pts_direct(K, obs_pts, obs_ns, src_pts, src_ns, vec, params, float_type):
    """out[i] = sum_j K(src_j - obs_i) vec_j over all source points (pairs at
    r2 == 0 skipped); the reference entry point tectosaur/farfield.py:15-38.
    Returns float64 [n_obs * 3]."""
    obs_pts = np.asarray(obs_pts, dtype=np.float64).reshape((-1, 3))
    src_pts = np.asarray(src_pts, dtype=np.float64).reshape((-1, 3))
    obs_ns = np.asarray(obs_ns, dtype=np.float64).reshape((-1, 3))
    src_ns = np.asarray(src_ns, dtype=np.float64).reshape((-1, 3))
    vec = np.asarray(vec, dtype=np.float64).reshape((-1, 3))
    po = morton_order(obs_pts)
    ps = morton_order(src_pts)
    prec = prec_of(float_type)
    eng = FarfieldEngine(K, params, obs_pts[po], obs_ns[po], src_pts[ps], src_ns[ps],
                         precisions=(prec,))
    dev = eng.device
    v = torch.as_tensor(np.ascontiguousarray(vec[ps]), dtype=torch.float64, device=dev)
    eng.set_strengths(v, prec)
    eng.run(prec)
    out = torch.empty((eng.n_obs, 3), dtype=torch.float64, device=dev)
    eng.reduce_pts(prec, out)
    res = np.empty((eng.n_obs, 3))
    res[po] = out.cpu().numpy()
    return res.reshape(-1)
