"""The dense pair matrix ``farfield_tris`` and ``RegularizedDenseIntegralOp``
(tectosaur/ops/dense_integral_op.py:11-84, tectosaur/assemble.cl:105-138),
and a ``CompositeOp`` of subset operators (the fault / surface blocks of
tectosaur/qd/model_helpers.py:22-39), on the device against the reference
fixture, the CPU oracle and the sparse operator."""
import os

import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.quadrature import gauss4d_tri

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden')
PARAMS = [1.0, 0.25]
NQ = (3, 3, 3, 2, 3)
THR = 2.0


def test_farfield_tris_golden():
    # tectosaur tests/test_integral_op.py:82-94: farfield_tris('elasticH3',
    # [1, 0.25], pts, [[0, 1, 2]], [[3, 4, 5]], 3, float32)
    from tectosaur_amd.dense_integral_op import farfield_tris
    pts = np.array([[1, 0, 0], [2, 0, 0], [1, 1, 0], [5, 0, 0], [6, 0, 0], [5, 1, 0]], float)
    out = farfield_tris('elasticH3', [1.0, 0.25], pts, np.array([[0, 1, 2]]),
                        np.array([[3, 4, 5]]), 3, np.float32)
    g = np.load(os.path.join(GOLD, 'test_farfield_two_tris.npy'))
    assert out.shape == g.shape == (1, 3, 3, 1, 3, 3)
    np.testing.assert_allclose(out, g, rtol=0, atol=2e-6 * np.abs(g).max())


@pytest.mark.parametrize('K', ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3',
                               'elasticRT3', 'elasticRA3', 'elasticRH3'])
def test_farfield_tris_vs_oracle(oracle, K):
    from tectosaur_amd.dense_integral_op import farfield_tris
    pts, tris = mesh.make_rect(5, 4, [[-1, 0, 1], [1, 0, 1], [1, 0.4, -1], [-1, 0.2, -1]])
    obs, src = tris[:10], tris[4:]
    got = farfield_tris(K, PARAMS, pts, obs, src, 2, np.float64)
    assert got.shape == (10, 3, 3, src.shape[0], 3, 3)
    q = gauss4d_tri(2, 2)
    # every (obs, src) pair, skipping the same triangle (coincident rule points)
    all_tris = np.concatenate([obs, src])
    pairs = np.array([(i, 10 + j) for i in range(10) for j in range(src.shape[0])
                      if not np.array_equal(obs[i], src[j])])
    want = oracle.pairs_integrate(K, q[0], q[1], pts, all_tris, pairs, False, PARAMS)
    for k, (i, j) in enumerate(pairs):
        blk = got[i, :, :, j - 10]
        np.testing.assert_allclose(blk, want[k], rtol=0, atol=1e-12 * np.abs(want[k]).max())


def _split_mesh():
    pts, tris, n_surf = mesh.fault_surface_mesh(9, 5, 3, surf_w=2000.0, fault_L=1000.0)
    return pts, tris, n_surf


@pytest.mark.parametrize('subsets', ['full', 'fault_from_surface', 'overlapping'])
def test_dense_op_equals_sparse_op_and_oracle(oracle, subsets):
    from tectosaur_amd import triangle_rules
    from tectosaur_amd.dense_integral_op import RegularizedDenseIntegralOp
    from tectosaur_amd.sparse_integral_op import RegularizedSparseIntegralOp
    pts, tris, n_surf = _split_mesh()
    n = tris.shape[0]
    obs, src = {'full': (np.arange(n), np.arange(n)),
                'fault_from_surface': (np.arange(n_surf, n), np.arange(n_surf)),
                'overlapping': (np.arange(40, 120), np.arange(0, 90))}[subsets]
    Kn, Kf = 'elasticRH3', 'elasticRH3'
    dense = RegularizedDenseIntegralOp(*NQ, THR, Kn, Kf, PARAMS, pts, tris, np.float64,
                                       obs_subset=obs, src_subset=src)
    sparse = RegularizedSparseIntegralOp(*NQ, THR, Kn, Kf, PARAMS, pts, tris, np.float64,
                                         obs_subset=obs, src_subset=src)
    assert dense.shape == sparse.shape == (9 * obs.shape[0], 9 * src.shape[0])
    v = np.random.default_rng(31).normal(size=dense.shape[1])
    want = sparse.dot(v)
    got = dense.dot(v)
    assert np.linalg.norm(got - want) <= 1e-11 * np.linalg.norm(want)
    # the oracle's dense matrix: uncorrected near blocks where nonzero, the
    # farfield rule elsewhere
    rules = {'co': triangle_rules.coincident_quad(NQ[0]),
             'ea': triangle_rules.edge_adj_quad(NQ[1]),
             'va': triangle_rules.vertex_adj_quad(NQ[2], NQ[2], NQ[2]),
             'near': gauss4d_tri(NQ[4], NQ[4]), 'far': gauss4d_tri(NQ[3], NQ[3])}
    blocks = oracle.nearfield_blocks(pts, tris, *NQ, THR, Kn, Kf, PARAMS, rules,
                                     obs_subset=obs, src_subset=src)
    Mn = oracle.nearfield_dense(blocks, obs.shape[0], src.shape[0], corrected=False)
    q = rules['far']
    all_tris = np.concatenate([tris[obs], tris[src]])
    pairs = np.array([(i, obs.shape[0] + j) for i in range(obs.shape[0])
                      for j in range(src.shape[0])])
    F = oracle.pairs_integrate(Kf, q[0], q[1], pts, all_tris, pairs, False, PARAMS)
    F = F.reshape((obs.shape[0], src.shape[0], 9, 9)).transpose(0, 2, 1, 3).reshape(Mn.shape)
    M = np.where(np.abs(Mn) > 0, Mn, F)
    M[np.isnan(M)] = 0.0
    assert np.isfinite(M).all() and np.isfinite(dense.mat).all()
    assert np.abs(dense.mat - M).max() <= 1e-12 * np.abs(M).max()


def test_composite_of_subset_ops_equals_full():
    """The fault / surface block decomposition of one operator: four subset
    operators in a CompositeOp equal the operator on the whole mesh, host
    and device products, FP64 farfield."""
    import torch
    from tectosaur_amd.combinators import CompositeOp
    from tectosaur_amd.sparse_integral_op import RegularizedSparseIntegralOp
    pts, tris, n_surf = _split_mesh()
    n = tris.shape[0]
    S, F = np.arange(n_surf), np.arange(n_surf, n)
    mk = lambda o, s: RegularizedSparseIntegralOp(*NQ, THR, 'elasticRH3', 'elasticRH3', PARAMS,
                                                  pts, tris, np.float64, obs_subset=o,
                                                  src_subset=s)
    comp = CompositeOp((mk(S, S), 0, 0), (mk(S, F), 0, 9 * n_surf),
                       (mk(F, S), 9 * n_surf, 0), (mk(F, F), 9 * n_surf, 9 * n_surf))
    full = mk(np.arange(n), np.arange(n))
    v = np.random.default_rng(32).normal(size=9 * n)
    want = full.dot(v)
    assert np.linalg.norm(comp.dot(v) - want) <= 1e-12 * np.linalg.norm(want)
    vd = torch.as_tensor(v, device='cuda')
    got = comp.dot_dev(vd, prec='f64').cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
