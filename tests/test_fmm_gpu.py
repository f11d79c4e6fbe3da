"""Treecode farfield (FMMFarfieldOp) against the direct farfield on the
device.  The reference's test (tests/fmm/test_taylor.py:12-71) compares its
treecode with TriToTriDirectFarfieldOp on two 20x20 plates (quadrature order
2, mac 2.5), coincident or 6 apart, for U3, RT3, RA3, RH3 at order 8 in
FP64, to 5 decimals.  The same comparison here at order 8, plus: the
error falls with the order, the FP32 treecode
matches the FP64 one to FP32 accuracy, and with a huge acceptance parameter
(every interaction direct) the treecode equals the direct sum."""
import numpy as np
import pytest

from tectosaur_amd import mesh

pytestmark = pytest.mark.gpu

KS = ['elasticU3', 'elasticRT3', 'elasticRA3', 'elasticRH3']
CORNERS = [[-1.0, -1.0, 0], [-1.0, 1.0, 0], [1.0, 1.0, 0], [1.0, -1.0, 0]]


def plates(offset, n=20):
    m_src = mesh.make_rect(n, n, CORNERS)
    m_obs = mesh.make_rect(n, n, CORNERS)
    m_obs[0][:, 0] += offset
    pts, tris = mesh.concat_two_no_remove_duplicates(m_src, m_obs)
    ns = m_src[1].shape[0]
    return pts, tris, np.arange(ns, tris.shape[0]), np.arange(ns)


def run(K, offset, order, float_type, mac=2.5, quad=2, seed=123987):
    from tectosaur_amd.fmm import FMMFarfieldOp
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris, obs, src = plates(offset)
    v = np.random.default_rng(seed).random(src.shape[0] * 9)
    direct = TriToTriDirectFarfieldOp(quad, K, [1.0, 0.25], pts, tris, np.float64, obs, src)
    y1 = direct.dot(v)
    op = FMMFarfieldOp(mac, 64, order)(quad, K, [1.0, 0.25], pts, tris, float_type, obs, src)
    y2 = op.dot(v)
    return y1, y2, op


@pytest.mark.parametrize('K', KS)
@pytest.mark.parametrize('offset', [0.0, 6.0])
def test_treecode_matches_direct_like_reference(K, offset):
    y1, y2, op = run(K, offset, 8, np.float64)
    rel = np.linalg.norm(y2 - y1) / np.linalg.norm(y1)
    print(K, offset, 'rel', rel, op.tree.stats)
    np.testing.assert_almost_equal(y1, y2, 5)
    assert rel < 2e-5
    if offset == 6.0:
        assert op.tree.stats['p2p_tiles'] == 0


@pytest.mark.parametrize('K', ['elasticU3', 'elasticRH3'])
def test_treecode_error_falls_with_order(K):
    errs = []
    for order in (2, 3, 4, 5, 6, 7, 8):
        y1, y2, op = run(K, 0.0, order, np.float64, mac=2.0)
        errs.append(np.linalg.norm(y2 - y1) / np.linalg.norm(y1))
    print(K, errs)
    assert all(errs[i + 1] < errs[i] for i in range(6))
    assert errs[4] < 0.05 * errs[0]


@pytest.mark.parametrize('K', KS)
def test_treecode_f32_matches_f64(K):
    y1, y64, _ = run(K, 0.0, 4, np.float64)
    _, y32, _ = run(K, 0.0, 4, np.float32)
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) < 1e-5


@pytest.mark.parametrize('K', KS)
@pytest.mark.parametrize('ft,tol', [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_treecode_all_direct_equals_direct(K, ft, tol):
    # mac so large that no node is accepted: every interaction is p2p
    y1, y2, op = run(K, 0.0, 2, ft, mac=1e6)
    assert op.tree.stats['m2p'] == 0
    assert np.linalg.norm(y2 - y1) / np.linalg.norm(y1) < tol


