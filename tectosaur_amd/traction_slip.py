"""Traction -> slip solves on planar faults (the request the server batches).

Mirrors the reference's caller get_traction_to_slip
(tectosaur/qd/full_model.py:179-197) with the operator recipe of
tectosaur/qd/model_helpers.py:22-39:

    rhs   = -M (t / sm)                     (P1 mass matrix, MassOp)
    cm    = continuity + free-edge constraints (build_constraint_matrix)
    solve cm^T H cm x = cm^T rhs,  slip = cm x

with H = RegularizedSparseIntegralOp(K_near = K_far = elasticRH3, params
[1, nu]) -- the reference's recipe (build_elastic_op uses the regularized
kernel for both halves).  Using the regularized kernel for the farfield too
keeps the discrete operator consistent: mixing RH3 near pairs with H3 far
pairs leaves integration-by-parts boundary terms that do not telescope, and
the constrained system's condition number then grows several times faster
with refinement (measured: 49 vs 7.8 at 506 triangles).  K_far = elasticH3
remains available.  The farfield is the FP32/FP64 N-body of far_p2p.hip;
the nearfield is FP64.

``continuity_cm`` builds the same cm as build_constraint_matrix applied to
continuity_constraints + free_edge_constraints (zero field) for a single
surface, directly from the vertex grouping (every DOF at a vertex equals
the smallest such DOF; DOFs at free-edge vertices are zero).
"""
import math

import attr
import numpy as np
import scipy.sparse
import torch

from . import _native as N
from . import mesh as mesh_mod
from .constraints import find_free_edges
from .mass_op import MassOp
from .sparse_integral_op import RegularizedSparseIntegralOp

DEFAULT_ORDERS = dict(nq_coincident=6, nq_edge_adj=6, nq_vert_adjacent=6, nq_far=1, nq_near=5,
                      near_threshold=2.5, quad_mass_order=3)


def continuity_cm(tris):
    """cm of continuity + zero free-edge constraints on one surface."""
    tris = np.asarray(tris, dtype=np.int64)
    n = tris.shape[0]
    n_dofs = 9 * n
    free = set()
    for tri_idx, edge_idx in find_free_edges(tris):
        free.add(int(tris[tri_idx, edge_idx]))
        free.add(int(tris[tri_idx, (edge_idx + 1) % 3]))
    vert = np.repeat(tris.reshape(-1), 3)            # vertex of each dof
    comp = np.tile(np.arange(3), 3 * n)              # component of each dof
    key = vert * 3 + comp
    dofs = np.arange(n_dofs)
    order = np.lexsort((dofs, key))
    skey = key[order]
    first = np.ones(n_dofs, dtype=bool)
    first[1:] = skey[1:] != skey[:-1]
    rep = np.maximum.accumulate(np.where(first, np.arange(n_dofs), 0))
    root = np.empty(n_dofs, dtype=np.int64)
    root[order] = order[rep]                         # smallest dof sharing (vertex, comp)
    is_free_vert = np.zeros(int(tris.max()) + 1, dtype=bool)
    if free:
        is_free_vert[np.fromiter(free, dtype=np.int64)] = True
    keep = ~is_free_vert[vert]
    independent = keep & (root == dofs)
    col_of = np.full(n_dofs, -1, dtype=np.int64)
    col_of[independent] = np.arange(int(independent.sum()))
    rows = dofs[keep]
    cols = col_of[root[keep]]
    return scipy.sparse.csr_matrix((np.ones(rows.shape[0]), (rows, cols)),
                                   shape=(n_dofs, int(independent.sum())))


@attr.s
class FaultRequest:
    """One served solve: a planar rectangular fault and a traction."""
    rid = attr.ib()
    strike = attr.ib()       # degrees
    dip = attr.ib()          # degrees
    length = attr.ib()       # m
    aspect = attr.ib()       # length / width
    n_tris = attr.ib()       # target triangle count
    traction = attr.ib()     # (3,) uniform part
    seed = attr.ib()         # seed of the N(0, 0.1) traction noise
    arrival = attr.ib(default=0.0)


def random_requests(n, rate, seed=0, n_min=512, n_max=32768):
    """The served workload: Poisson arrivals at ``rate`` per second; strike
    U(0, 360), dip U(20, 90), length U(5, 50) km, aspect U(1, 4), triangle
    count log-uniform in [n_min, n_max], traction U(-1, 1)^3 (+ noise)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for i in range(n):
        out.append(FaultRequest(
            rid=i, strike=rng.uniform(0, 360), dip=rng.uniform(20, 90),
            length=rng.uniform(5e3, 50e3), aspect=rng.uniform(1, 4),
            n_tris=int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))),
            traction=rng.uniform(-1, 1, 3), seed=int(rng.integers(1 << 31)), arrival=float(t[i])))
    return out


def fault_mesh(req):
    """Rectangle length x (length / aspect) in the fault plane, centred at
    the origin, meshed with nx x nz squares (2 nx nz ~ n_tris triangles)."""
    L = req.length
    W = L / req.aspect
    nz = max(1, int(round(math.sqrt(req.n_tris / (2.0 * req.aspect)))))
    nx = max(1, int(round(req.n_tris / (2.0 * nz))))
    th = math.radians(req.strike)
    dl = math.radians(req.dip)
    s = np.array([math.cos(th), math.sin(th), 0.0])
    d = np.array([-math.sin(th) * math.cos(dl), math.cos(th) * math.cos(dl), -math.sin(dl)])
    a, b = 0.5 * L * s, 0.5 * W * d
    corners = [a + b, -a + b, -a - b, a - b]
    return mesh_mod.make_rect(nx + 1, nz + 1, corners)


# treecode of the served 'tree' mode: expansion order, acceptance parameter
# and the smallest problem (triangles) that gets one -- below it the direct
# FP32 N-body is faster than the treecode's fixed per-matvec passes
DEFAULT_TREE = dict(order=3, mac=2.0, min_tris=3000)


def prepare_problem(pts, tris, traction, sm=3e10, orders=None, K_far='elasticRH3', tree=None):
    """The host-only half of a traction -> slip set-up (mesh-side work that
    needs no device): pair classification, farfield layout, mass matrix,
    constraint matrix and right-hand side, and (``tree`` = dict(order,
    mac)) the treecode's tree and interaction lists.  Safe to run on a
    worker thread while the device is busy with other solves."""
    from .nearfield_op import nearfield_plan
    from .sparse_farfield_op import farfield_plan
    from .kernels import kernels as KERNELS
    o = dict(DEFAULT_ORDERS)
    o.update(orders or {})
    pts = np.asarray(pts, dtype=np.float64)
    tris = np.asarray(tris, dtype=np.int64)
    idx = np.arange(tris.shape[0])
    K = KERNELS[K_far]
    M = MassOp(o['quad_mass_order'], pts, tris)
    t = np.asarray(traction, dtype=np.float64).reshape(-1)
    far_plan = farfield_plan(o['nq_far'], pts, tris, idx, idx, K.surf_curl_obs or K.surf_curl_src)
    out = dict(pts=pts, tris=tris, orders=o, K_far=K_far,
               near_plan=nearfield_plan(pts, tris, idx, idx, o['near_threshold']),
               far_plan=far_plan, M=M, cm=continuity_cm(tris), rhs=-M.dot(t / sm), tree=None)
    if tree is not None and tris.shape[0] >= tree.get('min_tris', 0):
        from .fmm import tree_plan
        out['tree'] = dict(tree, plan=tree_plan(far_plan['o_pts'], far_plan['s_pts'],
                                                tree['mac']))
    return out


class TractionSlipProblem:
    """Device operators and right-hand side of one traction -> slip solve."""

    def __init__(self, pts, tris, traction, sm=3e10, nu=0.25, orders=None,
                 farfield_precisions=('f32', 'f64'), K_far='elasticRH3', prepared=None):
        if prepared is None:
            prepared = prepare_problem(pts, tris, traction, sm, orders, K_far)
        o = prepared['orders']
        self.pts = prepared['pts']
        self.tris = prepared['tris']
        n = self.tris.shape[0]
        self.n_tris = n
        self.H = RegularizedSparseIntegralOp(
            o['nq_coincident'], o['nq_edge_adj'], o['nq_vert_adjacent'], o['nq_far'],
            o['nq_near'], o['near_threshold'], 'elasticRH3', prepared['K_far'], [1.0, nu],
            self.pts, self.tris, np.float32, farfield_precisions=farfield_precisions,
            near_plan=prepared['near_plan'], far_plan=prepared['far_plan'],
            keep_uncorrected=False)
        self.device = self.H.device
        if 'f32' in farfield_precisions:
            # inner (FP32 / treecode) matvecs read FP32 nearfield blocks
            self.H.nearfield.use_f32_inner()
        if prepared.get('tree') is not None:
            tr = prepared['tree']
            self.H.farfield.attach_tree(order=tr['order'], mac=tr['mac'], plan=tr['plan'])
        self.M = prepared['M']
        self.cm = prepared['cm']
        self.rhs = N.upload(prepared['rhs'], torch.float64, self.device)

    @property
    def n_dofs(self):
        return self.cm.shape[1]

    @property
    def pair_evals(self):
        return self.H.farfield.pair_evals


def request_traction(req, n_tris):
    rng = np.random.default_rng(req.seed)
    t = np.tile(np.asarray(req.traction, dtype=np.float64), n_tris * 3)
    return t + rng.normal(0.0, 0.1, t.shape[0])
