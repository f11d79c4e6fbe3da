"""Parity of the HIP nearfield path (pair integrator, nearfield operator,
BSR matvec) with the CPU oracle.  Bar: FP64 elementwise agreement,
|gpu - oracle| <= 1e-10 * max|block| for every 9x9 block."""
import os

import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.find_near_adj import find_close_or_touching, split_adjacent_close
from tectosaur_amd.quadrature import gauss4d_tri
from tectosaur_amd import triangle_rules

pytestmark = pytest.mark.gpu
ALLK = ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3', 'elasticRT3', 'elasticRA3',
        'elasticRH3']
GOLD = os.path.join(os.path.dirname(__file__), 'golden')
PARAMS = [3e10, 0.25]


def blockwise_close(a, b, tol=1e-10):
    a = a.reshape((-1, 81))
    b = b.reshape((-1, 81))
    scale = np.abs(b).max(axis=1, keepdims=True)
    scale[scale == 0] = 1.0
    err = np.abs(a - b) / scale
    return err.max() <= tol, err.max()


def small_mesh(flip_some=False):
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 1)
    pts = pts + 0.02 * np.random.default_rng(11).normal(size=pts.shape)
    if flip_some:
        tris = tris.copy()
        tris[::4] = tris[::4][:, [1, 0, 2]]
    return pts, tris


def classes(pts, tris, thr=2.0):
    from tectosaur_amd.nearfield_op import resolve_ea_rotation
    close = find_close_or_touching(pts, tris, pts, tris, thr)
    near, va4, ea6 = split_adjacent_close(close, tris, tris)
    va = np.concatenate([va4, np.zeros((va4.shape[0], 1), np.int64)], axis=1)
    ea = resolve_ea_rotation(tris, ea6)
    co = np.stack([np.arange(tris.shape[0])] * 2, axis=1)
    return co, ea, va, near


@pytest.mark.parametrize('K', ALLK)
@pytest.mark.parametrize('flip', [False, True])
def test_pair_classes_vs_oracle(oracle, K, flip):
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts, tris = small_mesh(flip)
    co, ea, va, near = classes(pts, tris)
    if flip:
        assert ea[:, 4].any()  # the flipped mesh exercises src_flip
    pi = PairsIntegrator(K, PARAMS, np.float64, 2, 3, pts, tris)
    cases = [('co', co[:12], triangle_rules.coincident_quad(3), True, pi.coincident(3, co[:12])),
             ('ea', ea[:24], triangle_rules.edge_adj_quad(3), True, pi.edge_adj(3, ea[:24])),
             ('va', va[:24], triangle_rules.vertex_adj_quad(3, 3, 3), True,
              pi.vert_adj(3, va[:24])),
             ('near', near[:24], gauss4d_tri(3, 3), False, pi.nearfield(near[:24])),
             ('corr', near[:24], gauss4d_tri(2, 2), False, pi.correction(near[:24], False)),
             ('corr0', co[:12], gauss4d_tri(2, 2), True, pi.correction(co[:12], True))]
    for name, prs, q, check0, got in cases:
        want = oracle.pairs_integrate(K, q[0], q[1], pts, tris, prs, check0, PARAMS)
        ok, err = blockwise_close(got, want)
        assert ok, (K, name, err)


@pytest.mark.parametrize('lanes', [1, 4, 16, 64])
@pytest.mark.parametrize('K', ['elasticH3', 'elasticRH3', 'elasticRT3'])
def test_lanes_per_pair_agree(oracle, lanes, K):
    import torch
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts, tris = small_mesh(True)
    co, ea, va, near = classes(pts, tris)
    pi = PairsIntegrator(K, PARAMS, np.float64, 2, 3, pts, tris)
    q = pi.quad_to_gpu(triangle_rules.edge_adj_quad(2))
    p = torch.as_tensor(ea[:37].astype(np.int32), device='cuda')
    got = pi.pairs_quad_dev(q, p, True, lanes_per_pair=lanes).cpu().numpy()
    qh = triangle_rules.edge_adj_quad(2)
    want = oracle.pairs_integrate(K, qh[0], qh[1], pts, tris, ea[:37], True, PARAMS)
    ok, err = blockwise_close(got, want)
    assert ok, err


@pytest.mark.parametrize('chunks', [1, 2, 5, 11, 12, 16, 44, 48, 88])
def test_rh_rule_chunks_vs_oracle(oracle, chunks):
    # RH3 one lane per pair and rule chunk (lanes_per_pair = -chunks): the
    # chunks' sums added in order reproduce the integrals, for the singular
    # classes (flipped sources included) and a rule that does not split
    # evenly; beyond 11 chunks four (up to 44) or eight waves of a workgroup
    # take as many chunks of the same pairs and add them in LDS (pair counts
    # not multiples of 64)
    import torch
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts, tris = small_mesh(True)
    co, ea, va, near = classes(pts, tris)
    pi = PairsIntegrator('elasticRH3', PARAMS, np.float64, 2, 3, pts, tris)
    for prs, q in ((co[:17], triangle_rules.coincident_quad(3)),
                   (ea[:37], triangle_rules.edge_adj_quad(3)),
                   (va[:29], triangle_rules.vertex_adj_quad(3, 3, 3))):
        qd = pi.quad_to_gpu(q)
        p = torch.as_tensor(prs.astype(np.int32), device='cuda')
        got = pi.pairs_quad_dev(qd, p, True, lanes_per_pair=-chunks).cpu().numpy()
        want = oracle.pairs_integrate('elasticRH3', q[0], q[1], pts, tris, prs, True, PARAMS)
        ok, err = blockwise_close(got, want)
        assert ok, (chunks, prs.shape, err)


@pytest.mark.parametrize('K', ['elasticH3', 'elasticRH3'])
def test_r0_test_modes_agree_on_rules_without_equal_points(K):
    # the singular rules have no point of equal obs and src parameters, so
    # testing r2 == 0 alone (mode 2, what the integrator asks for with these
    # rules) gives the bits of the full test (mode 1), for every lane count
    import torch
    from tectosaur_amd import _native as N
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts, tris = small_mesh(True)
    co, ea, va, near = classes(pts, tris)
    pi = PairsIntegrator(K, PARAMS, np.float64, 2, 3, pts, tris)
    for prs, q in ((co, triangle_rules.coincident_quad(3)), (ea, triangle_rules.edge_adj_quad(3)),
                   (va, triangle_rules.vertex_adj_quad(3, 3, 3))):
        qd = pi.quad_to_gpu(q)
        assert not qd.has_equal_params
        p = torch.as_tensor(prs.astype(np.int32), device='cuda')
        for lanes in (1, 4, 16, 64):
            outs = []
            for mode in (1, 2):
                o = torch.empty((prs.shape[0], 81), dtype=torch.float64, device='cuda')
                N.call('tct_pairs_integrate', pi.kid, qd[0].shape[0], N.ptr(qd[0]),
                       N.ptr(qd[1]), N.ptr(pi.gpu_pts), N.ptr(pi.gpu_tris), prs.shape[0],
                       N.ptr(p), p.shape[1], mode, pi.params.ctypes.data, lanes, N.ptr(o),
                       N.stream_ptr())
                outs.append(o)
            assert torch.equal(outs[0], outs[1]), (K, lanes)


def test_golden_two_tris_on_device():
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts = np.array([[1, 0, 0], [2, 0, 0], [1, 1, 0], [5, 0, 0], [6, 0, 0], [5, 1, 0]], float)
    tris = np.array([[0, 1, 2], [3, 4, 5]])
    pi = PairsIntegrator('elasticH3', [1.0, 0.25], np.float32, 3, 3, pts, tris)
    out = pi.nearfield(np.array([[0, 1]]))
    g = np.load(os.path.join(GOLD, 'test_farfield_two_tris.npy'))
    np.testing.assert_allclose(out.reshape(g.shape), g, rtol=0, atol=2e-6 * np.abs(g).max())


def test_golden_vert_adjacent_on_device():
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, -1, 0], [2, 0, 0]], float)
    tris = np.array([[1, 2, 0], [1, 3, 4]])
    pi = PairsIntegrator('elasticH3', [1.0, 0.25], np.float32, 1, 1, pts, tris)
    out = pi.vert_adj(3, np.array([[0, 1, 0, 0]]))
    g = np.load(os.path.join(GOLD, 'test_gpu_vert_adjacent.npy'))
    np.testing.assert_allclose(out, g, rtol=0, atol=1e-6 * np.abs(g).max())


def oracle_rules(nq_co, nq_ea, nq_va, nq_far, nq_near):
    return {'co': triangle_rules.coincident_quad(nq_co), 'ea': triangle_rules.edge_adj_quad(nq_ea),
            'va': triangle_rules.vertex_adj_quad(nq_va, nq_va, nq_va),
            'near': gauss4d_tri(nq_near, nq_near), 'far': gauss4d_tri(nq_far, nq_far)}


@pytest.mark.parametrize('Kn,Kf', [('elasticRH3', 'elasticH3'), ('elasticRT3', 'elasticT3'),
                                   ('elasticU3', 'elasticU3'), ('elasticRA3', 'elasticA3')])
def test_nearfield_op_vs_oracle(oracle, Kn, Kf):
    from tectosaur_amd.nearfield_op import RegularizedNearfieldIntegralOp
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    n = tris.shape[0]
    nq = (4, 4, 3, 2, 3)
    op = RegularizedNearfieldIntegralOp(pts, tris, np.arange(n), np.arange(n), nq[0], nq[1],
                                        nq[2], nq[3], nq[4], 2.0, Kn, Kf, PARAMS, np.float32,
                                        keep_classes=True)
    ref = oracle.nearfield_blocks(pts, tris, *nq, 2.0, Kn, Kf, PARAMS, oracle_rules(*nq))
    rows, cols, data = [], [], []
    for name in ('co', 'ea', 'va', 'near'):
        prs, dofs, acc, corr = op.classes[name]
        rprs, rdofs, rcorr, racc = ref[name]
        assert np.array_equal(prs, rprs), name
        ok, err = blockwise_close(acc.cpu().numpy(), racc)
        assert ok, (name, 'uncorrected', err)
        ok, err = blockwise_close((acc - corr).cpu().numpy(), rcorr, 1e-9)
        assert ok, (name, 'corrected', err)
        rows.append(dofs[:, 0]); cols.append(dofs[:, 1]); data.append(rcorr)
    v = np.random.default_rng(12).normal(size=n * 9)
    want = oracle.bcoo_matvec(np.concatenate(rows), np.concatenate(cols),
                              np.concatenate(data), v, op.shape)
    got = op.dot(v)
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


def test_bsr_matvec_vs_dense():
    import torch
    from tectosaur_amd.sparse import BCOOMatrix
    rng = np.random.default_rng(13)
    nb = 40
    rows = rng.integers(0, nb, 300)
    cols = rng.integers(0, nb, 300)
    key = np.unique(rows * nb + cols)
    rows, cols = key // nb, key % nb
    data = rng.normal(size=(rows.shape[0], 9, 9))
    A = np.zeros((9 * nb, 9 * nb))
    for r, c, d in zip(rows, cols, data):
        A[9 * r:9 * r + 9, 9 * c:9 * c + 9] = d
    M = BCOOMatrix(rows, cols, data, A.shape)
    x = rng.normal(size=9 * nb)
    np.testing.assert_allclose(M.dot(x), A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(M.to_dense(), A)
    # FP32 block copy (inner products of mixed-precision solves): the blocks
    # rounded to FP32, FP64 accumulation -- equal to the dense product with
    # FP32-rounded entries, to FP64 roundoff
    B = M.to_bsr().make_f32()
    xd = torch.as_tensor(x, device='cuda')
    y32 = B.dot_dev(xd, prec='f32').cpu().numpy()
    A32 = A.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(y32, A32 @ x, rtol=1e-12, atol=1e-12 * np.abs(A32 @ x).max())
    assert np.linalg.norm(y32 - A @ x) <= 1e-6 * np.linalg.norm(A @ x)
    # FP64 products are unchanged by the copy
    np.testing.assert_array_equal(B.dot_dev(xd, prec='f64').cpu().numpy(), M.dot(x))


def test_bsr_empty_rows_and_matrix():
    from tectosaur_amd.sparse import BCOOMatrix
    M = BCOOMatrix(np.array([3]), np.array([0]), np.ones((1, 9, 9)), (45, 45))
    out = M.dot(np.ones(45))
    assert np.all(out[:27] == 0) and np.all(out[27:36] == 9) and np.all(out[36:] == 0)
    E = BCOOMatrix(np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros((0, 9, 9)), (18, 18))
    assert np.all(E.dot(np.ones(18)) == 0)


@pytest.mark.parametrize('Kn,Kf', [('elasticRH3', 'elasticH3'), ('elasticRT3', 'elasticT3')])
def test_icosphere_l5_sampled_pairs(oracle, Kn, Kf):
    # icosphere level 5 (20,480 triangles), all rules n = 8, near threshold 3:
    # every class assembled on the device; a seeded sample of pairs of each
    # class is integrated by the oracle and compared elementwise, both the
    # accurate blocks and the corrected ones (minus the K_far farfield rule),
    # and the BSR product of the corrected blocks is compared on sampled
    # block rows with the oracle's blocks of those rows.
    from tectosaur_amd.nearfield_op import RegularizedNearfieldIntegralOp
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 5)
    n = tris.shape[0]
    assert n == 20480
    nq = (8, 8, 8, 2, 8)
    op = RegularizedNearfieldIntegralOp(pts, tris, np.arange(n), np.arange(n), *nq, 3.0, Kn, Kf,
                                        PARAMS, np.float64, keep_classes=True)
    rules = oracle_rules(*nq)
    rng = np.random.default_rng(14)
    for name in ('co', 'ea', 'va', 'near'):
        prs, dofs, acc, corr = op.classes[name]
        assert prs.shape[0] > 0
        sel = np.sort(rng.choice(prs.shape[0], size=min(16, prs.shape[0]), replace=False))
        check0 = name != 'near'
        want = oracle.pairs_integrate(Kn, rules[name][0], rules[name][1], pts, tris, prs[sel],
                                      check0, PARAMS)
        a = acc.cpu().numpy()
        got = a[sel]
        ok, err = blockwise_close(got, want)
        assert ok, (name, err)
        assert np.isfinite(a).all()
        # corrected blocks: the error is measured against the accurate block's
        # size (the difference of the two rules cancels for near pairs)
        wcorr = oracle.pairs_integrate(Kf, rules['far'][0], rules['far'][1], pts, tris,
                                       prs[sel, :2], name == 'co', PARAMS)
        gc = (acc - corr).cpu().numpy()[sel].reshape((-1, 81))
        wc = (want - wcorr).reshape((-1, 81))
        scale = np.abs(want.reshape((-1, 81))).max(axis=1)
        assert (np.abs(gc - wc).max(axis=1) <= 1e-10 * scale).all(), name
    # BSR product on sampled block rows
    v = rng.normal(size=9 * n)
    y = op.dot(v)
    rows = np.sort(rng.choice(n, size=8, replace=False))
    for r in rows:
        yr = np.zeros(9)
        for name in ('co', 'ea', 'va', 'near'):
            prs, dofs, _, _ = op.classes[name]
            mine = np.nonzero(dofs[:, 0] == r)[0]
            if mine.size == 0:
                continue
            check0 = name != 'near'
            B = oracle.pairs_integrate(Kn, rules[name][0], rules[name][1], pts, tris,
                                       prs[mine], check0, PARAMS)
            B = B - oracle.pairs_integrate(Kf, rules['far'][0], rules['far'][1], pts, tris,
                                           prs[mine, :2], name == 'co', PARAMS)
            B = B.reshape((-1, 9, 9))
            for k, j in enumerate(dofs[mine, 1]):
                yr += B[k] @ v[9 * j:9 * j + 9]
        assert np.linalg.norm(y[9 * r:9 * r + 9] - yr) <= 1e-10 * np.linalg.norm(yr), r


@pytest.mark.parametrize('accumulate', [False, True])
def test_bsr_long_rows_vs_dense(accumulate):
    """Block rows of 0, 1, 62, 63, 64, 126, 127 and 200 blocks: the kernel
    stages the column indices of 63 blocks at a time, so rows longer than
    63 run several batches.  FP64 blocks and the FP32 copy, with and
    without accumulation, against the dense product."""
    import torch
    from tectosaur_amd.sparse import BCOOMatrix
    rng = np.random.default_rng(17)
    lens = [0, 1, 62, 63, 64, 126, 127, 200, 5]
    nbc = 220
    rows, cols = [], []
    for r, L in enumerate(lens):
        c = np.sort(rng.choice(nbc, size=L, replace=False))
        rows.append(np.full(L, r))
        cols.append(c)
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    data = rng.normal(size=(rows.shape[0], 9, 9))
    shape = (9 * len(lens), 9 * nbc)
    A = np.zeros(shape)
    for r, c, d in zip(rows, cols, data):
        A[9 * r:9 * r + 9, 9 * c:9 * c + 9] = d
    x = rng.normal(size=shape[1])
    y0 = rng.normal(size=shape[0])
    B = BCOOMatrix(rows, cols, data, shape).to_bsr().make_f32()
    xd = torch.as_tensor(x, device='cuda')
    A32 = A.astype(np.float32).astype(np.float64)
    for prec, M in (('f64', A), ('f32', A32)):
        out = torch.as_tensor(y0.copy(), device='cuda')
        B.dot_dev(xd, out, accumulate=accumulate, prec=prec)
        want = (y0 if accumulate else 0) + M @ x
        np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-12,
                                   atol=1e-12 * np.abs(want).max())
