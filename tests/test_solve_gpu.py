"""Traction -> slip solves on the device: converged residual checked under
the FP64 oracle operator, mixed vs FP64 agreement, and the server's
alone-vs-in-the-mix identity and failure reporting."""
import numpy as np
import pytest
import torch

from tectosaur_amd.quadrature import gauss2d_tri, gauss4d_tri
from tectosaur_amd import triangle_rules
from tectosaur_amd.traction_slip import (FaultRequest, TractionSlipProblem, fault_mesh,
                                         request_traction)

pytestmark = pytest.mark.gpu
ORD = dict(nq_coincident=4, nq_edge_adj=4, nq_vert_adjacent=3, nq_far=1, nq_near=3,
           near_threshold=2.5, quad_mass_order=3)


def oracle_apply(oracle, pts, tris, u, nu=0.25, K_far='elasticRH3'):
    o = ORD
    rules = {'co': triangle_rules.coincident_quad(o['nq_coincident']),
             'ea': triangle_rules.edge_adj_quad(o['nq_edge_adj']),
             'va': triangle_rules.vertex_adj_quad(*(o['nq_vert_adjacent'],) * 3),
             'near': gauss4d_tri(o['nq_near'], o['nq_near']),
             'far': gauss4d_tri(o['nq_far'], o['nq_far'])}
    blocks = oracle.nearfield_blocks(pts, tris, o['nq_coincident'], o['nq_edge_adj'],
                                     o['nq_vert_adjacent'], o['nq_far'], o['nq_near'],
                                     o['near_threshold'], 'elasticRH3', K_far, [1.0, nu],
                                     rules)
    y = np.zeros(tris.shape[0] * 9)
    for prs, _, corr, _ in blocks.values():
        y += oracle.bcoo_matvec(prs[:, 0], prs[:, 1], corr, u, (y.shape[0], y.shape[0]))
    qx, qw = gauss2d_tri(o['nq_far'])
    y += oracle.farfield_tris(K_far, qx, qw, pts, tris, tris, u, [1.0, nu])
    return y


def small_problem(n_tris=200, seed=5):
    req = FaultRequest(0, 35.0, 55.0, 8000.0, 1.5, n_tris, np.array([0.3, -0.7, 0.2]), seed)
    pts, tris = fault_mesh(req)
    return pts, tris, request_traction(req, tris.shape[0])


@pytest.mark.parametrize('K_far', ['elasticRH3', 'elasticH3'])
@pytest.mark.parametrize('mode', ['f64', 'mixed'])
def test_solve_residual_under_oracle_operator(oracle, mode, K_far):
    from tectosaur_amd.solver import solve_constrained
    pts, tris, t = small_problem()
    prob = TractionSlipProblem(pts, tris, t, orders=ORD, K_far=K_far)
    u, res = solve_constrained(prob.H, prob.cm, prob.rhs, tol=1e-8, mode=mode, restart=60)
    assert res.converged and res.rel_residual <= 1e-8
    u = u.cpu().numpy()
    b = prob.cm.T @ prob.rhs.cpu().numpy()
    r = b - prob.cm.T @ oracle_apply(oracle, pts, tris, u, K_far=K_far)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.1e-8
    if mode == 'mixed':
        assert res.n_matvec['f32'] > res.n_matvec['f64']


@pytest.mark.parametrize('order,mac', [(3, 2.0), (4, 2.5)])
def test_tree_solve_residual_under_oracle_operator(oracle, order, mac):
    # FP32 treecode farfield in the Arnoldi cycles, exact FP64 residuals:
    # the converged residual is the exact operator's, checked with the oracle.
    from tectosaur_amd.solver import solve_constrained
    from tectosaur_amd.traction_slip import prepare_problem
    pts, tris, t = small_problem(600, 9)
    prep = prepare_problem(pts, tris, t, orders=ORD, tree=dict(order=order, mac=mac))
    prob = TractionSlipProblem(None, None, None, prepared=prep)
    assert prob.H.farfield.tree.stats['m2p'] > 0
    u, res = solve_constrained(prob.H, prob.cm, prob.rhs, tol=1e-8, mode='tree', restart=100)
    assert res.converged and res.rel_residual <= 1e-8
    assert res.n_matvec['tree'] > res.n_matvec['f64'] > 0
    u = u.cpu().numpy()
    b = prob.cm.T @ prob.rhs.cpu().numpy()
    r = b - prob.cm.T @ oracle_apply(oracle, pts, tris, u)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.1e-8


def test_mixed_matches_f64_solution():
    from tectosaur_amd.solver import solve_constrained
    pts, tris, t = small_problem(300, 7)
    prob = TractionSlipProblem(pts, tris, t, orders=ORD)
    u64, r64 = solve_constrained(prob.H, prob.cm, prob.rhs, tol=1e-10, mode='f64', restart=80)
    u32, r32 = solve_constrained(prob.H, prob.cm, prob.rhs, tol=1e-10, mode='mixed', restart=80)
    assert r64.converged and r32.converged
    u64 = u64.cpu().numpy(); u32 = u32.cpu().numpy()
    assert np.linalg.norm(u32 - u64) <= 1e-8 * np.linalg.norm(u64)


def _mix_requests(n=8, seed=3):
    """n requests spanning 512 to 8,192 triangles (log-uniform), the first
    and last pinned to the ends of the range."""
    from tectosaur_amd.traction_slip import random_requests
    reqs = random_requests(n, rate=0.0, seed=seed, n_min=512, n_max=8192)
    reqs[0].n_tris, reqs[-1].n_tris = 512, 8192
    return reqs


@pytest.mark.parametrize('batch_nbody', [False, True])
def test_server_result_same_alone_and_in_mix(batch_nbody):
    # batch_nbody: the solves' FP32 N-bodies run as one segment-table launch
    # per tick; a solve alone runs the same path (a batch of one)
    from tectosaur_amd.serve import Server
    reqs = _mix_requests()
    mix = Server(mode='mixed', max_active=8, restart=500, orders=ORD, keep_slip=True,
                 batch_nbody=batch_nbody)
    recs_mix = mix.run(reqs, realtime=False)
    assert mix.max_concurrent >= 4  # solves were in flight together
    if batch_nbody:
        assert mix.nbody_launches > 0
    for req, rm in zip(reqs, recs_mix):
        alone = Server(mode='mixed', max_active=1, restart=500, orders=ORD, keep_slip=True,
                       batch_nbody=batch_nbody)
        ra = alone.run([req], realtime=False)[0]
        assert rm.status == ra.status == 'converged' and rm.rel_residual <= 1e-8
        assert rm.iterations == ra.iterations
        # identical kernels on identical buffers: agreement to roundoff (bitwise in practice)
        assert np.abs(rm.slip - ra.slip).max() <= 1e-12 * np.abs(ra.slip).max()


def test_tree_server_result_same_alone_and_in_mix():
    from tectosaur_amd.serve import Server
    from tectosaur_amd.traction_slip import random_requests
    reqs = _mix_requests(seed=8)
    # the treecode above 2,000 triangles, the FP32 direct farfield below
    tree = dict(order=4, mac=2.5, min_tris=2000)
    # two builder threads (each with its own build stream; captures serialised)
    mix = Server(mode='tree', max_active=8, restart=500, orders=ORD, keep_slip=True, tree=tree,
                 builders=2, batch_nbody=True)
    recs_mix = mix.run(reqs, realtime=False)
    for req, rm in zip(reqs, recs_mix):
        ra = Server(mode='tree', max_active=1, restart=500, orders=ORD, keep_slip=True,
                    tree=tree, batch_nbody=True).run([req], realtime=False)[0]
        assert rm.status == ra.status == 'converged' and rm.rel_residual <= 1e-8
        assert rm.iterations == ra.iterations
        assert np.abs(rm.slip - ra.slip).max() <= 1e-12 * np.abs(ra.slip).max()


def test_server_reports_failures():
    from tectosaur_amd.serve import Server
    from tectosaur_amd.traction_slip import random_requests
    reqs = random_requests(2, rate=0.0, seed=4, n_min=256, n_max=300)
    s = Server(mode='mixed', max_active=2, restart=30, maxiter=1, tol=1e-14, orders=ORD)
    recs = s.run(reqs, realtime=False)
    assert all(r.status == 'maxiter' for r in recs)
    bad = random_requests(1, rate=0.0, seed=5, n_min=256, n_max=300)
    bad[0].traction = np.array([np.nan, 0.0, 0.0])
    r = Server(mode='mixed', max_active=1, orders=ORD).run(bad, realtime=False)[0]
    assert r.status == 'nonfinite'
