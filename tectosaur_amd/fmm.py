"""Treecode farfield: drop-in for the reference's ``FMMFarfieldOp``
(tectosaur/ops/sparse_farfield_op.py:73-118, engine tectosaur/fmm/tsfmm.py).

The reference builds an octree over triangles and runs p2m, m2m, m2p and p2p
passes (a treecode: no local expansions) with spherical-harmonic expansions
of the elastic kernels U3, RT3, RA3 and RH3.  This build keeps the operator
and its parameters (``mac``, ``pts_per_cell``, ``order``) and re-designs the
rest for the GPU (far_tree.h):

  * the tree is a binary tree over the 64-point Morton tiles that the direct
    N-body already packs (a node is a contiguous range of tiles), so the
    p2p pass reuses the direct kernel's source records and per-tile origins;
  * one interaction list per observation tile (one wavefront), built on the
    host with the reference's acceptance test
    ``mac * R_src + R_obs < 0.98 * separation`` (tectosaur/fmm/
    traversal.cpp:124-160); accepted nodes are evaluated from multipoles
    (m2p), the rest refined down to source tiles evaluated directly (p2p);
  * multipoles are Cartesian Taylor moments of 1/r with per-kernel Laplace
    charges; moments of parents come from children (m2m), one launch per
    level up to the highest level the lists use.

``pts_per_cell`` sets the smallest source node evaluated directly: ranges of
``max(64, pts_per_cell)`` points rounded up to a power-of-two number of
tiles.  The observation side always works per 64-point tile.
"""
import numpy as np
import torch

from . import _native as N
from .farfield import GROUP
from .kernels import kernels as KERNELS
from .sparse_farfield_op import TriToTriDirectFarfieldOp, farfield_plan

TREE_KERNELS = ('elasticU3', 'elasticRT3', 'elasticRA3', 'elasticRH3')
EVAL_ORDERS = (2, 3, 4, 5, 6, 7, 8)
N_CHANNELS = {'elasticU3': 4, 'elasticRA3': 4, 'elasticRT3': 13, 'elasticRH3': 12}


def n_terms(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def _tile_bounds(pts, n_pad, lo_tiles, hi_tiles, box=None):
    """Centre (bounding-box midpoint) and radius of point ranges given in
    tiles [lo, hi); padding points beyond pts.shape[0] are excluded.  The
    ranges must be consecutive and cover [lo[0], hi[-1]).  ``box`` (bmin,
    bmax) may give the ranges' bounding boxes when already known."""
    n = pts.shape[0]
    lo = np.minimum(lo_tiles * GROUP, n)
    hi = np.minimum(hi_tiles * GROUP, n)
    cnt = hi - lo
    if np.any(cnt <= 0):
        raise ValueError('empty tree node')
    if box is None:
        box = (np.minimum.reduceat(pts, lo, axis=0), np.maximum.reduceat(pts, lo, axis=0))
    bmin, bmax = box
    c = 0.5 * (bmin + bmax)
    owner = np.repeat(np.arange(lo.shape[0]), cnt)
    diff = pts[lo[0]:hi[-1]] - c[owner]
    d2 = np.einsum('ij,ij->i', diff, diff)
    r = np.sqrt(np.maximum.reduceat(d2, lo - lo[0]))
    return c, r, box


class SourceTree:
    """Binary tree over the tiles of a packed source point set.

    Nodes are numbered level by level from the tiles up: node t < n_tiles is
    tile t; level l has ceil(n_tiles / 2^l) nodes, node i of level l covers
    tiles [i 2^l, min((i + 1) 2^l, n_tiles)) and has children 2i and 2i + 1
    of level l - 1 (the second when it exists)."""

    def __init__(self, pts, n_pad):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        self.n_tiles = n_pad // GROUP
        ext = float(np.max(np.ptp(pts, axis=0))) if pts.shape[0] > 1 else 1.0
        rmin = 1e-9 * (ext if ext > 0 else 1.0)
        counts = [self.n_tiles]
        while counts[-1] > 1:
            counts.append((counts[-1] + 1) // 2)
        self.levels = []
        cr, lo_all, hi_all, lev_all = [], [], [], []
        off = 0
        box = None
        for lvl, cnt in enumerate(counts):
            span = 1 << lvl
            lo = np.arange(cnt) * span
            hi = np.minimum(lo + span, self.n_tiles)
            if box is not None:
                # a parent's box is the union of its children's (exact)
                bmin, bmax = box
                if bmin.shape[0] % 2:
                    bmin = np.concatenate([bmin, bmin[-1:]])
                    bmax = np.concatenate([bmax, bmax[-1:]])
                box = (np.minimum(bmin[0::2], bmin[1::2]), np.maximum(bmax[0::2], bmax[1::2]))
            c, r, box = _tile_bounds(pts, n_pad, lo, hi, box)
            cr.append(np.concatenate([c, np.maximum(r, rmin)[:, None]], axis=1))
            lo_all.append(lo)
            hi_all.append(hi)
            lev_all.append(np.full(cnt, lvl))
            self.levels.append((off, cnt))
            off += cnt
        self.n_nodes = off
        self.node_cr = np.ascontiguousarray(np.concatenate(cr))
        self.tile_lo = np.concatenate(lo_all)
        self.tile_hi = np.concatenate(hi_all)
        self.level = np.concatenate(lev_all)
        self.root = self.n_nodes - 1
        # first child and child count of every node (tiles have none)
        self.child0 = np.full(self.n_nodes, -1, dtype=np.int64)
        self.nchild = np.zeros(self.n_nodes, dtype=np.int64)
        for lvl in range(1, len(self.levels)):
            p_off, n_par = self.levels[lvl]
            c_off, n_child = self.levels[lvl - 1]
            i = np.arange(n_par)
            self.child0[p_off + i] = c_off + 2 * i
            self.nchild[p_off + i] = np.where(2 * i + 1 < n_child, 2, 1)


def interaction_lists(obs_cr, tree, mac, leaf_level=0):
    """Per-observation-tile lists: (list_start [n_obs+1], list) int32, list
    entries >= 0 for m2p nodes and -(t + 1) for p2p source tiles t.
    obs_cr [n_obs_tiles, 4]: centre and radius of each observation tile."""
    n_obs = obs_cr.shape[0]
    po = np.arange(n_obs)
    pn = np.full(n_obs, tree.root)
    m_o, m_n, p_o, p_t = [], [], [], []
    nc = tree.node_cr
    while po.size:
        sep = np.linalg.norm(obs_cr[po, :3] - nc[pn, :3], axis=1)
        acc = mac * nc[pn, 3] + obs_cr[po, 3] < 0.98 * sep
        m_o.append(po[acc])
        m_n.append(pn[acc])
        po, pn = po[~acc], pn[~acc]
        leaf = tree.level[pn] <= leaf_level
        if np.any(leaf):
            lo = tree.tile_lo[pn[leaf]]
            cnt = tree.tile_hi[pn[leaf]] - lo
            p_o.append(np.repeat(po[leaf], cnt))
            start = np.repeat(lo - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
            p_t.append(start + np.arange(int(cnt.sum())))
        po, pn = po[~leaf], pn[~leaf]
        c0 = tree.child0[pn]
        two = tree.nchild[pn] == 2
        po = np.concatenate([po, po[two]])
        pn = np.concatenate([c0, c0[two] + 1])
    obs = np.concatenate(m_o + p_o)
    code = np.concatenate(m_n + [-(t + 1) for t in p_t]).astype(np.int64)
    order = np.argsort(obs, kind='stable')
    counts = np.bincount(obs, minlength=n_obs)
    start = np.zeros(n_obs + 1, dtype=np.int64)
    start[1:] = np.cumsum(counts)
    if start[-1] >= 2 ** 31:
        raise N.TctError('interaction list too long for int32 indices')
    stats = dict(m2p=int(sum(a.size for a in m_n)), p2p_tiles=int(sum(a.size for a in p_t)))
    return start.astype(np.int32), code[order].astype(np.int32), stats


def obs_tile_cr(pts, n_pad):
    n_tiles = n_pad // GROUP
    lo = np.arange(n_tiles)
    c, r, _ = _tile_bounds(np.ascontiguousarray(pts, dtype=np.float64), n_pad, lo, lo + 1)
    return np.concatenate([c, r[:, None]], axis=1)


def segments(start, n_waves):
    """Split every tile's list [start[t], start[t+1]) into segments of at
    most S entries, S chosen so that about ``n_waves`` segments result
    (bounded to [2, 64]).  Returns (seg [n_seg, 3] int32 = (tile, lo, hi),
    seg_off [n_tiles + 1] int32)."""
    start = np.asarray(start, dtype=np.int64)
    lens = np.diff(start)
    S = int(min(64, max(2, -(-int(start[-1]) // max(n_waves, 1)))))
    nseg = np.maximum(1, -(-lens // S))
    seg_off = np.zeros(lens.shape[0] + 1, dtype=np.int64)
    seg_off[1:] = np.cumsum(nseg)
    tile = np.repeat(np.arange(lens.shape[0]), nseg)
    k = np.arange(int(seg_off[-1])) - seg_off[tile]
    lo = start[tile] + k * S
    hi = np.minimum(lo + S, start[tile + 1])
    seg = np.stack([tile, lo, hi], axis=1).astype(np.int32)
    return np.ascontiguousarray(seg), seg_off.astype(np.int32), S


def tree_plan(obs_pts, src_pts, mac=3.0, pts_per_cell=64):
    """Host half of the treecode set-up (no device work): source tree,
    observation tile bounds and interaction lists for point sets already in
    packing order."""
    from .farfield import round_up
    tree = SourceTree(src_pts, round_up(src_pts.shape[0], GROUP))
    obs_cr = obs_tile_cr(obs_pts, round_up(obs_pts.shape[0], GROUP))
    leaf_level = max(0, int(np.ceil(np.log2(max(pts_per_cell, GROUP) / GROUP))))
    start, lst, stats = interaction_lists(obs_cr, tree, float(mac), leaf_level)
    return dict(tree=tree, obs_cr=obs_cr, start=start, list=lst, stats=stats, mac=float(mac))


class TreeFarfield:
    """Treecode evaluation of a FarfieldEngine's kernel (same packed layouts,
    same partial-sum output)."""

    def __init__(self, engine, obs_pts=None, src_pts=None, order=4, mac=3.0, pts_per_cell=64,
                 plan=None, n_waves=None):
        if engine.K.name not in TREE_KERNELS:
            raise NotImplementedError('%s has no treecode form (U3, RT3, RA3, RH3 do)'
                                      % engine.K.name)
        if order not in EVAL_ORDERS:
            raise ValueError('treecode order must be one of %s' % (EVAL_ORDERS,))
        if plan is None:
            plan = tree_plan(obs_pts, src_pts, mac, pts_per_cell)
        self.engine = engine
        self.order = order
        self.mac = plan['mac']
        self.tree = plan['tree']
        self.obs_cr = plan['obs_cr']
        self.stats = plan['stats']
        start, lst = plan['start'], plan['list']
        if (self.tree.n_tiles * GROUP != engine.n_src_pad or
                self.obs_cr.shape[0] * GROUP != engine.n_obs_pad):
            raise ValueError('treecode plan does not match the engine layout')
        dev = engine.device
        self.n_obs_tiles = engine.n_obs_pad // GROUP
        self.list = N.upload(lst, None, dev)
        self.node_cr = N.upload(self.tree.node_cr, torch.float64, dev)
        self.C = N_CHANNELS[engine.K.name]
        self.T = n_terms(order)
        if n_waves is None:
            # enough wavefronts to hide the latency of each wave's serial
            # list: 64 per CU of an MI355X (256 CUs), fixed rather than read
            # from the device since the segmentation sets the summation order
            n_waves = 64 * 256
        seg, seg_off, self.seg_len = segments(start, n_waves)
        self.n_seg = seg.shape[0]
        self.seg = N.upload(seg, None, dev)
        self.seg_off = N.upload(seg_off, None, dev)
        self.chunks = 1  # the reduced output has one chunk
        # upward pass: one launch per level, only up to the highest level any
        # list refers to.  Three one-launch forms measured slower at 4k-32k
        # triangles: the block that finishes a parent's last child going on
        # to the parent (device-scope fences between levels, 1.4x), one
        # workgroup forming the small top levels node after node (3x at 4k),
        # and one workgroup forming them with every node of a level at once
        # (180 us for the top five levels at 8k against 55 us per level: four
        # waves cannot hide the translation's LDS and FP64 latency)
        used = lst[lst >= 0]
        self.max_level = int(self.tree.level[used].max()) if used.size else 0
        self._moments = {}
        self._partial = None
        self._seg_partial = None

    def moments(self, prec):
        if prec not in self._moments:
            dt = torch.float32 if prec == 'f32' else torch.float64
            self._moments[prec] = torch.empty(self.tree.n_nodes * self.C * self.T, dtype=dt,
                                              device=self.engine.device)
        return self._moments[prec]

    @property
    def partial(self):
        if self._partial is None:
            self._partial = torch.empty(self.engine.n_out * self.engine.n_obs_pad,
                                        dtype=torch.float64, device=self.engine.device)
            self._seg_partial = torch.empty(self.n_seg * self.engine.n_out * GROUP,
                                            dtype=torch.float64, device=self.engine.device)
        return self._partial

    def run(self, prec='f32'):
        """Moments, upward pass and evaluation from the engine's current
        source records (set by its gather); returns the partial sums."""
        e = self.engine
        pid = e._prec_id(prec)
        st = N.stream_ptr()
        mom = self.moments(prec)
        rec = e.rec(prec)
        N.call('tct_tree_p2m', e.kid, pid, self.order, e.n_src, self.tree.n_tiles, N.ptr(rec),
               N.ptr(e.src_org), N.ptr(self.node_cr), N.ptr(mom), st)
        for lvl in range(1, self.max_level + 1):
            p_off, n_par = self.tree.levels[lvl]
            c_off, n_child = self.tree.levels[lvl - 1]
            N.call('tct_tree_m2m', e.kid, pid, self.order, p_off, n_par, c_off, n_child,
                   N.ptr(self.node_cr), N.ptr(mom), st)
        obs = e.loc32 if prec == 'f32' else e.obs_soa64
        out = self.partial
        N.call('tct_tree_eval', e.kid, pid, self.order, e.n_obs_pad, N.ptr(obs), N.ptr(e.obs_org),
               N.ptr(rec), N.ptr(e.src_org), N.ptr(self.node_cr), N.ptr(mom), self.n_seg,
               N.ptr(self.seg), N.ptr(self.list), e.params.ctypes.data, float(e.eps2),
               N.ptr(self._seg_partial), st)
        N.call('tct_tree_reduce', e.kid, self.n_obs_tiles, e.n_obs_pad, N.ptr(self.seg_off),
               N.ptr(self._seg_partial), N.ptr(out), st)
        return out

    def pair_evals_equivalent(self):
        """Direct pair evaluations done by the p2p pass (64 x 64 per tile pair)."""
        return self.stats['p2p_tiles'] * GROUP * GROUP


class FMMFarfieldOp:
    """Factory with the reference's signature: FMMFarfieldOp(mac,
    pts_per_cell, order)(nq_far, K_name, params, pts, tris, float_type,
    obs_subset, src_subset)."""

    def __init__(self, mac, pts_per_cell, order):
        self.mac = mac
        self.pts_per_cell = pts_per_cell
        self.order = order

    def __call__(self, nq_far, K_name, params, pts, tris, float_type, obs_subset, src_subset):
        return FMMFarfieldOpImpl(nq_far, K_name, params, pts, tris, float_type, obs_subset,
                                 src_subset, self.mac, self.pts_per_cell, self.order)


class FMMFarfieldOpImpl(TriToTriDirectFarfieldOp):
    """Treecode tri-to-tri farfield: the direct operator's gather and
    scatter around a treecode evaluation.  ``float_type`` selects FP32 or
    FP64 arithmetic; ``dot_dev(..., prec='f64')`` also evaluates directly
    in FP64 (the exact operator) for residuals."""

    def __init__(self, nq_far, K_name, params, pts, tris, float_type, obs_subset, src_subset,
                 mac, pts_per_cell, order, plan=None, precisions=None):
        if K_name not in TREE_KERNELS:
            raise NotImplementedError('%s has no treecode form (U3, RT3, RA3, RH3 do)' % K_name)
        K = KERNELS[K_name]
        obs_subset = np.asarray(obs_subset, dtype=np.int64)
        src_subset = np.asarray(src_subset, dtype=np.int64)
        if plan is None:
            plan = farfield_plan(nq_far, pts, tris, obs_subset, src_subset,
                                 K.surf_curl_obs or K.surf_curl_src)
        super().__init__(nq_far, K_name, params, pts, tris, float_type, obs_subset, src_subset,
                         precisions=precisions, plan=plan)
        self.attach_tree(order=order, mac=mac, pts_per_cell=pts_per_cell)

    def dot_dev(self, v_dev, out_dev=None, prec=None):
        prec = prec or self.prec
        if prec == 'tree':
            prec = 'f32'
        return self.tree_dot_dev(v_dev, out_dev, prec)

    def nearfield_dot(self, v):
        return self.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.dot(v)

    def farfield_dot(self, v):
        return self.dot(v)
