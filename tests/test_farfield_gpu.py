"""Parity of the HIP farfield path (libtct_hip.so) with the CPU oracle.

Tolerances: FP64 path relative L2 <= 1e-12 (same formula, different
summation order); FP32 path relative L2 <= 1e-5 (the bar the farfield
workloads state for FP32 evaluation with FP64 accumulation)."""
import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.quadrature import gauss2d_tri

pytestmark = pytest.mark.gpu
KS = ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3']
TOL = {'f32': 1e-5, 'f64': 1e-12}


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def random_pts(n, seed):
    rng = np.random.default_rng(seed)
    obs = rng.random((n, 3))
    ns = rng.random((n, 3))
    ns /= np.linalg.norm(ns, axis=1)[:, None]
    src = rng.random((n, 3))
    w = rng.random((n, 3)).reshape(-1)
    return obs, ns, src, ns.copy(), w


@pytest.mark.parametrize('prec', ['f32', 'f64'])
@pytest.mark.parametrize('K', KS)
def test_farfield_pts_direct(oracle, K, prec):
    # the reference's N-body smoke case (tests/test_farfield.py:53-92): n = 1000 random points
    from tectosaur_amd.farfield import farfield_pts_direct
    obs, on, src, sn, w = random_pts(1000, 100)
    ft = np.float32 if prec == 'f32' else np.float64
    out = farfield_pts_direct(K, obs, on, src, sn, w, [1.0, 0.25], ft)
    ref = oracle.farfield_pts(K, obs, on, src, sn, w, [1.0, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


@pytest.mark.parametrize('n_obs,n_src', [(1, 1), (1, 200), (63, 65), (130, 7)])
@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_ragged_sizes(oracle, n_obs, n_src, prec):
    from tectosaur_amd.farfield import farfield_pts_direct
    rng = np.random.default_rng(n_obs * 1000 + n_src)
    obs = rng.random((n_obs, 3)); src = rng.random((n_src, 3))
    on = rng.normal(size=(n_obs, 3)); on /= np.linalg.norm(on, axis=1)[:, None]
    sn = rng.normal(size=(n_src, 3)); sn /= np.linalg.norm(sn, axis=1)[:, None]
    w = rng.normal(size=n_src * 3)
    ft = np.float32 if prec == 'f32' else np.float64
    out = farfield_pts_direct('elasticH3', obs, on, src, sn, w, [3e10, 0.25], ft)
    ref = oracle.farfield_pts('elasticH3', obs, on, src, sn, w, [3e10, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


def test_coincident_points_skipped(oracle):
    # r2 == 0 pairs are skipped (farfield_direct.cl:89-91): obs set == src set
    from tectosaur_amd.farfield import farfield_pts_direct
    obs, on, _, _, w = random_pts(300, 7)
    for prec, ft in (('f32', np.float32), ('f64', np.float64)):
        out = farfield_pts_direct('elasticU3', obs, on, obs, on, w, [1.0, 0.25], ft)
        ref = oracle.farfield_pts('elasticU3', obs, on, obs, on, w, [1.0, 0.25])
        assert np.all(np.isfinite(out))
        assert rel_l2(out, ref) <= TOL[prec]


def two_surfaces(n=8, sep=2.0, w=1.0):
    # tectosaur tests/test_farfield.py:10-25
    m1 = mesh.make_rect(n, n, [[-w, 0, w], [-w, 0, -w], [w, 0, -w], [w, 0, w]])
    m2 = mesh.make_rect(n, n, [[-w, sep, w], [-w, sep, -w], [w, sep, -w], [w, sep, w]])
    m = mesh.concat(m1, m2)
    s1 = np.arange(m1[1].shape[0])
    s2 = s1[-1] + 1 + s1
    return m, s1, s2


@pytest.mark.parametrize('prec', ['f32', 'f64'])
@pytest.mark.parametrize('K', KS)
@pytest.mark.parametrize('nq', [1, 2])
def test_tri_tri_separated(oracle, K, prec, nq):
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    (pts, tris), s1, s2 = two_surfaces()
    ft = np.float32 if prec == 'f32' else np.float64
    op = TriToTriDirectFarfieldOp(nq, K, [1.0, 0.25], pts, tris, ft, s1, s2)
    v = np.random.default_rng(1).random(op.shape[1])
    out = op.dot(v)
    qx, qw = gauss2d_tri(nq)
    ref = oracle.farfield_tris(K, qx, qw, pts, tris[s1], tris[s2], v, [1.0, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


@pytest.mark.parametrize('prec', ['f32', 'f64'])
@pytest.mark.parametrize('K', KS)
def test_tri_tri_self(oracle, K, prec):
    # obs == src: the coincident quadrature-point pairs are skipped
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = mesh.make_rect(12, 9, [[-1, 0, 1], [-1, 0, -1], [1, 0, -1], [1, 0, 1]])
    n = tris.shape[0]
    ft = np.float32 if prec == 'f32' else np.float64
    op = TriToTriDirectFarfieldOp(2, K, [3e10, 0.25], pts, tris, ft, np.arange(n), np.arange(n))
    v = np.random.default_rng(2).normal(size=op.shape[1])
    out = op.dot(v)
    qx, qw = gauss2d_tri(2)
    ref = oracle.farfield_tris(K, qx, qw, pts, tris, tris, v, [3e10, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


def test_tri_tri_overlapping_subsets(oracle):
    # obs and src subsets overlap but differ: shared triangles are packed into
    # different groups, and their self pairs must still be skipped in FP32.
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = mesh.make_rect(10, 10, [[-1, 0, 1], [-1, 0, -1], [1, 0, -1], [1, 0, 1]])
    n = tris.shape[0]
    obs = np.arange(n // 2 + 20)
    src = np.arange(n // 2 - 20, n)
    for prec, ft in (('f32', np.float32), ('f64', np.float64)):
        op = TriToTriDirectFarfieldOp(2, 'elasticH3', [1.0, 0.25], pts, tris, ft, obs, src)
        v = np.random.default_rng(3).normal(size=op.shape[1])
        out = op.dot(v)
        qx, qw = gauss2d_tri(2)
        ref = oracle.farfield_tris('elasticH3', qx, qw, pts, tris[obs], tris[src], v, [1.0, 0.25])
        assert rel_l2(out, ref) <= TOL[prec]


def fault_2k_mesh():
    # vertical fault 2 km x 2 km, 32 x 32 squares -> 2048 triangles
    return mesh.make_rect(33, 33, [[-1000, 0, 0], [1000, 0, 0], [1000, 0, -2000],
                                   [-1000, 0, -2000]])


@pytest.mark.parametrize('prec', ['f32', 'f64'])
def test_fault_2k_u_farfield(oracle, prec):
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    ft = np.float32 if prec == 'f32' else np.float64
    op = TriToTriDirectFarfieldOp(1, 'elasticU3', [3e10, 0.25], pts, tris, ft, np.arange(n),
                                  np.arange(n))
    v = np.random.default_rng(0).standard_normal(18432)
    out = op.dot(v)
    qx, qw = gauss2d_tri(1)
    ref = oracle.farfield_tris('elasticU3', qx, qw, pts, tris, tris, v, [3e10, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


@pytest.mark.parametrize('cfg', [(64, 1, 1, 0), (128, 2, 3, 0), (256, 4, 7, 0), (256, 1, 32, 0),
                                 (256, 1, 8, 2), (128, 2, 16, 2), (64, 4, 24, 2), (64, 1, 24, 6),
                                 (256, 2, 5, 4), (128, 4, 3, 5)])
def test_tile_shapes_agree(cfg):
    # every launch shape computes the same sum; FP64 results agree to
    # roundoff; the XCD chunk order (flags 2) gives the bits of the plain
    # order with the same chunks
    from tectosaur_amd import _native as N
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    v = np.random.default_rng(9).normal(size=n * 9)
    base = TriToTriDirectFarfieldOp(1, 'elasticH3', [3e10, 0.25], pts, tris, np.float64,
                                    np.arange(n), np.arange(n)).dot(v)
    for ft, tol in ((np.float64, 1e-13), (np.float32, 1e-5)):
        op = TriToTriDirectFarfieldOp(1, 'elasticH3', [3e10, 0.25], pts, tris, ft, np.arange(n),
                                      np.arange(n), tile_cfg=N.TileCfg(*cfg))
        out = op.dot(v)
        assert rel_l2(out, base) <= tol
        if cfg[3] & 2:
            plain = TriToTriDirectFarfieldOp(1, 'elasticH3', [3e10, 0.25], pts, tris, ft,
                                             np.arange(n), np.arange(n),
                                             tile_cfg=N.TileCfg(cfg[0], cfg[1], cfg[2], 0))
            assert np.array_equal(plain.dot(v), out)


def test_xcd_order_needs_chunks_multiple_of_8():
    from tectosaur_amd import _native as N
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    op = TriToTriDirectFarfieldOp(1, 'elasticH3', [3e10, 0.25], pts, tris, np.float32,
                                  np.arange(n), np.arange(n), tile_cfg=N.TileCfg(256, 1, 12, 2))
    with pytest.raises(N.TctError, match='multiple of 8'):
        op.dot(np.ones(n * 9))


def test_deterministic():
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    op = TriToTriDirectFarfieldOp(1, 'elasticH3', [3e10, 0.25], pts, tris, np.float32,
                                  np.arange(n), np.arange(n))
    v = np.random.default_rng(10).normal(size=n * 9)
    a = op.dot(v)
    b = op.dot(v)
    assert np.array_equal(a, b)


CURLK = ['elasticRT3', 'elasticRA3', 'elasticRH3']


@pytest.mark.parametrize('prec', ['f32', 'f64'])
@pytest.mark.parametrize('K', CURLK)
@pytest.mark.parametrize('nq', [1, 2])
def test_surface_curl_tri_tri_separated(oracle, K, prec, nq):
    # matrix_free.cl farfield_tris_to_tris{K} with the vector codes of kernels.cl:100-206
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    (pts, tris), s1, s2 = two_surfaces()
    ft = np.float32 if prec == 'f32' else np.float64
    op = TriToTriDirectFarfieldOp(nq, K, [1.0, 0.25], pts, tris, ft, s1, s2)
    v = np.random.default_rng(21).random(op.shape[1])
    out = op.dot(v)
    qx, qw = gauss2d_tri(nq)
    ref = oracle.farfield_tris(K, qx, qw, pts, tris[s1], tris[s2], v, [1.0, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


@pytest.mark.parametrize('prec', ['f32', 'f64'])
@pytest.mark.parametrize('K', CURLK)
def test_surface_curl_tri_tri_self(oracle, K, prec):
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = mesh.make_rect(10, 8, [[-1, 0, 1], [-1, 0, -1], [1, 0, -1], [1, 0, 1]])
    pts = pts + 0.05 * np.random.default_rng(22).normal(size=pts.shape)
    n = tris.shape[0]
    ft = np.float32 if prec == 'f32' else np.float64
    op = TriToTriDirectFarfieldOp(2, K, [3e10, 0.25], pts, tris, ft, np.arange(n), np.arange(n))
    v = np.random.default_rng(23).normal(size=op.shape[1])
    out = op.dot(v)
    qx, qw = gauss2d_tri(2)
    ref = oracle.farfield_tris(K, qx, qw, pts, tris, tris, v, [3e10, 0.25])
    assert rel_l2(out, ref) <= TOL[prec]


def test_point_form_refuses_surface_curl_kernels():
    from tectosaur_amd.farfield import farfield_pts_direct
    from tectosaur_amd import _native as N
    p = np.random.default_rng(0).random((10, 3))
    with pytest.raises(N.TctError):
        farfield_pts_direct('elasticRH3', p, p, p, p, p.reshape(-1), [1.0, 0.25], np.float32)


@pytest.mark.parametrize('K', ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3', 'elasticRT3',
                               'elasticRA3', 'elasticRH3'])
def test_tiles_without_the_r2_test_change_nothing(K):
    # source tiles far from a wave's observation groups skip the r2 <= eps2
    # test; forcing the test everywhere (flags bit 0) must give the same
    # bits, for one point set packed twice and for overlapping subsets
    # packed differently (shared points in unrelated groups); so must the
    # records staged through LDS (bit 2), with and without the test
    from tectosaur_amd import _native as N
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    for obs, src in ((np.arange(n), np.arange(n)), (np.arange(n // 2 + 300),
                                                    np.arange(n // 2 - 300, n))):
        outs = []
        for flags in (0, 1, 4, 5):
            op = TriToTriDirectFarfieldOp(2, K, [3e10, 0.25], pts, tris, np.float32, obs, src,
                                          tile_cfg=N.TileCfg(128, 1, 4, flags))
            v = np.random.default_rng(4).normal(size=op.shape[1])
            outs.append(op.dot(v))
        assert np.all(np.isfinite(outs[0]))
        for o in outs[1:]:
            assert np.array_equal(outs[0], o), (K, obs.shape[0], src.shape[0])


@pytest.mark.parametrize('K', ['elasticH3', 'elasticRH3'])
def test_tile_skip_is_safe_for_any_eps2(K):
    # the tiles that skip the r2 test are chosen from the group extents the
    # caller passes, not from eps2, so an eps2 far below the packing bound
    # still tests every tile that can hold a pair at r2 <= eps2: the same
    # bits as the test on every tile, and finite (a point paired with itself
    # in one packing has r2 == 0 exactly)
    from tectosaur_amd import _native as N
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    pts, tris = fault_2k_mesh()
    n = tris.shape[0]
    outs = []
    for flags in (0, 1):
        op = TriToTriDirectFarfieldOp(2, K, [3e10, 0.25], pts, tris, np.float32, np.arange(n),
                                      np.arange(n), tile_cfg=N.TileCfg(128, 1, 4, flags))
        op.engine.eps2 = 1e-30
        v = np.random.default_rng(4).normal(size=op.shape[1])
        outs.append(op.dot(v))
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize('K', ['elasticH3', 'elasticRH3', 'elasticU3'])
def test_batched_nbody_equals_single_launches(K):
    # problems of different sizes (and subsets) in one segment-table launch:
    # each one's partial sums are bit for bit those of its own launch with
    # the same source chunks, also when the table spans several launches
    import torch
    from tectosaur_amd import _native as N
    from tectosaur_amd.farfield import run_batched
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    ops = []
    for k, nx in enumerate((5, 9, 17, 12, 7) * 6):
        pts, tris = mesh.make_rect(nx, nx + 2, [[-1, 0, 1], [1, 0, 1], [1, 0.3, -1],
                                                [-1, -0.2, -1]])
        n = tris.shape[0]
        obs = np.arange(n) if k % 3 else np.arange(n // 2, n)
        op = TriToTriDirectFarfieldOp(1, K, [3e10, 0.25], pts, tris, np.float32, obs,
                                      np.arange(n))
        v = torch.as_tensor(np.random.default_rng(k).normal(size=op.shape[1]), device='cuda')
        op._gather(v, 'f32')
        ops.append(op)
    single = []
    for op in ops:
        single.append(op.engine.run('f32').clone())
    outs = run_batched([op.engine for op in ops])
    torch.cuda.synchronize()
    assert len(outs) == len(ops) > 24
    for a, b in zip(single, outs):
        assert torch.equal(a, b)
    # a launch shape other than the engines' own gives the same bits
    outs2 = [o.clone() for o in run_batched([op.engine for op in ops], block=128,
                                            obs_per_lane=2)]
    for a, b in zip(single, outs2):
        assert torch.equal(a, b)
