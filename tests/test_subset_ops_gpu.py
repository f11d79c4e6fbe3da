"""Parity of the operators on observation / source triangle subsets with the
CPU oracle: RegularizedNearfieldIntegralOp (tectosaur/nearfield/
nearfield_op.py:76-168, the subset maps of :24-42) and the summed matvec
RegularizedSparseIntegralOp.dot / dot_dev (tectosaur/ops/
sparse_integral_op.py:45-55) on the fault / surface split that
tectosaur/qd/model_helpers.py:22-39 builds, on overlapping, disjoint and
unsorted subsets, and on the full set.

Bars: FP64 farfield and nearfield within 1e-12 relative L2 of the oracle;
FP32 farfield within 1e-5 relative L2."""
import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd import triangle_rules
from tectosaur_amd.quadrature import gauss2d_tri, gauss4d_tri

pytestmark = pytest.mark.gpu
PARAMS = [1.0, 0.25]
NQ = (3, 3, 3, 2, 3)   # coincident, edge-adjacent, vertex-adjacent, far, near
THR = 2.0


def fault_surface():
    pts, tris, n_surf = mesh.fault_surface_mesh(9, 5, 3, surf_w=2000.0, fault_L=1000.0)
    return pts, tris, n_surf


def subset_cases():
    pts, tris, n_surf = fault_surface()
    n = tris.shape[0]
    rng = np.random.default_rng(21)
    fault = np.arange(n_surf, n)
    surf = np.arange(n_surf)
    return (pts, tris), {
        'full': (np.arange(n), np.arange(n)),
        'fault_from_surface': (fault, surf),
        'surface_from_fault': (surf, fault),
        'fault_from_all': (fault, np.arange(n)),
        'overlapping': (np.arange(40, 120), np.arange(0, 90)),
        'unsorted': (rng.permutation(np.arange(20, 110)), rng.permutation(np.arange(50, n))),
    }


MESH, CASES = subset_cases()


def oracle_rules():
    nq_co, nq_ea, nq_va, nq_far, nq_near = NQ
    return {'co': triangle_rules.coincident_quad(nq_co),
            'ea': triangle_rules.edge_adj_quad(nq_ea),
            'va': triangle_rules.vertex_adj_quad(nq_va, nq_va, nq_va),
            'near': gauss4d_tri(nq_near, nq_near), 'far': gauss4d_tri(nq_far, nq_far)}


def oracle_near(oracle, obs, src, Kn, Kf):
    pts, tris = MESH
    blocks = oracle.nearfield_blocks(pts, tris, *NQ, THR, Kn, Kf, PARAMS, oracle_rules(),
                                     obs_subset=obs, src_subset=src)
    return (oracle.nearfield_dense(blocks, obs.shape[0], src.shape[0], True),
            oracle.nearfield_dense(blocks, obs.shape[0], src.shape[0], False), blocks)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize('case', sorted(CASES))
def test_nearfield_op_subsets_vs_oracle(oracle, case):
    from tectosaur_amd.nearfield_op import RegularizedNearfieldIntegralOp
    pts, tris = MESH
    obs, src = CASES[case]
    Kn, Kf = 'elasticRH3', 'elasticH3'
    Mc, Mu, blocks = oracle_near(oracle, obs, src, Kn, Kf)
    op = RegularizedNearfieldIntegralOp(pts, tris, obs, src, *NQ, THR, Kn, Kf, PARAMS,
                                        np.float64)
    assert op.shape == (obs.shape[0] * 9, src.shape[0] * 9)
    assert op.counts['co'] == np.intersect1d(obs, src).shape[0]
    for name, (prs, dofs, _, _) in blocks.items():
        assert op.counts[name] == prs.shape[0], name
    v = np.random.default_rng(22).normal(size=src.shape[0] * 9)
    assert rel(op.dot(v), Mc @ v) <= 1e-12
    assert rel(op.nearfield_no_correction_dot(v), Mu @ v) <= 1e-12
    D = op.to_dense()
    assert np.abs(D - Mc).max() <= 1e-12 * np.abs(Mc).max()


@pytest.mark.parametrize('case', sorted(CASES))
@pytest.mark.parametrize('Kn,Kf', [('elasticRH3', 'elasticRH3'), ('elasticRH3', 'elasticH3'),
                                   ('elasticU3', 'elasticU3'), ('elasticRT3', 'elasticT3'),
                                   ('elasticRA3', 'elasticA3')])
def test_summed_matvec_subsets_vs_oracle(oracle, case, Kn, Kf):
    """RegularizedSparseIntegralOp.dot (host arrays) and .dot_dev (device
    tensors), FP64 and FP32 farfield, against oracle farfield_tris over the
    subsets plus the oracle nearfield."""
    import torch
    from tectosaur_amd.sparse_integral_op import RegularizedSparseIntegralOp
    pts, tris = MESH
    obs, src = CASES[case]
    Mc, _, _ = oracle_near(oracle, obs, src, Kn, Kf)
    v = np.random.default_rng(23).normal(size=src.shape[0] * 9)
    qx, qw = gauss2d_tri(NQ[3])
    far = oracle.farfield_tris(Kf, qx, qw, pts, tris[obs], tris[src], v, PARAMS)
    want = far + Mc @ v
    for ft, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        op = RegularizedSparseIntegralOp(*NQ, THR, Kn, Kf, PARAMS, pts, tris, ft,
                                         obs_subset=obs, src_subset=src)
        assert op.shape == (obs.shape[0] * 9, src.shape[0] * 9)
        got = op.dot(v)
        assert rel(got, want) <= tol, (ft, rel(got, want))
        assert rel(op.farfield_dot(v), far) <= tol
        assert rel(op.nearfield_dot(v), Mc @ v) <= 1e-12
        vd = torch.as_tensor(v, device='cuda')
        out = torch.full((op.shape[0],), np.nan, dtype=torch.float64, device='cuda')
        op.dot_dev(vd, out)
        assert rel(out.cpu().numpy(), want) <= tol
