"""The farfield at the full sizes of the two large workloads, against the
CPU oracle (tectosaur/matrix_free.cl:19-107 restated) on seeded row
samples.  Rows are independent, so a sample's relative L2 error is exact
for those rows.

  * two-surface mesh (147,456 triangles, 3-point rule, 442,368 points):
    U/T/A/H, FP32 (<= 1e-5) and FP64 (<= 1e-12), 128 sampled rows;
  * fault + topography mesh (655,360 triangles, 3-point rule, 1,966,080
    points): direct FP32 H and T (the workload's farfield) and RH (the
    solve's farfield), <= 1e-5, 32 sampled rows.
"""
import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.quadrature import gauss2d_tri

pytestmark = pytest.mark.gpu
PARAMS = [3e10, 0.25]


def _rows_check(oracle, K, pts, tris, v, outs, n_rows, seed):
    n = tris.shape[0]
    rows = np.sort(np.random.default_rng(seed).choice(n, size=n_rows, replace=False))
    sel = (rows[:, None] * 9 + np.arange(9)[None, :]).reshape(-1)
    qx, qw = gauss2d_tri(1)
    ref = oracle.farfield_tris(K, qx, qw, pts, tris, tris, v, PARAMS, rows=rows)
    return {p: float(np.linalg.norm(y[sel] - ref) / np.linalg.norm(ref)) for p, y in outs.items()}


@pytest.fixture(scope='module')
def two_surface():
    pts, tris = mesh.two_surface_mesh()
    assert tris.shape[0] == 147456
    return pts, tris


@pytest.mark.parametrize('K', ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3'])
def test_two_surface_farfield_vs_oracle(oracle, two_surface, K):
    import torch
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = two_surface
    n = tris.shape[0]
    idx = np.arange(n)
    op = TriToTriDirectFarfieldOp(1, K, PARAMS, pts, tris, np.float32, idx, idx,
                                  precisions=('f32', 'f64'))
    v = np.random.default_rng(0).standard_normal(9 * n)
    vd = torch.as_tensor(v, device='cuda')
    outs = {p: op.dot_dev(vd, prec=p).cpu().numpy() for p in ('f32', 'f64')}
    err = _rows_check(oracle, K, pts, tris, v, outs, 128, 1)
    print(K, err, flush=True)
    assert err['f32'] <= 1e-5 and err['f64'] <= 1e-12, err


@pytest.fixture(scope='module')
def fault_topography():
    pts, tris = mesh.fault_topography_mesh()
    assert tris.shape[0] == 655360
    return pts, tris


@pytest.mark.parametrize('K', ['elasticH3', 'elasticT3', 'elasticRH3'])
def test_fault_topography_direct_f32_vs_oracle(oracle, fault_topography, K):
    import torch
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_topography
    n = tris.shape[0]
    idx = np.arange(n)
    op = TriToTriDirectFarfieldOp(1, K, PARAMS, pts, tris, np.float32, idx, idx,
                                  precisions=('f32',))
    v = np.random.default_rng(2).standard_normal(9 * n)
    vd = torch.as_tensor(v, device='cuda')
    outs = {'f32': op.dot_dev(vd, prec='f32').cpu().numpy()}
    del op
    torch.cuda.empty_cache()
    err = _rows_check(oracle, K, pts, tris, v, outs, 32, 3)
    print(K, err, flush=True)
    assert err['f32'] <= 1e-5, err
