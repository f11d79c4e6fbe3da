"""Interior evaluation on the device (interior.hip + the point farfield)
against the oracle, the reference's test_interior fixture and the
rigid-body identity."""
import os

import numpy as np
import pytest
import torch

from tectosaur_amd import mesh
from tectosaur_amd.quadrature import gauss2d_tri, gaussxw, map_to

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden')
KIDS = {'elasticU3': 0, 'elasticT3': 1, 'elasticA3': 2, 'elasticH3': 3, 'elasticRT3': 4}


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device='cuda')


def _case(seed=0):
    rng = np.random.default_rng(seed)
    pts, tris = mesh.make_rect(5, 4, [[0, 0, 0], [2, 0, 0.3], [2, 1.5, 0.1], [0, 1.5, 0]])
    n = 40
    obs = np.concatenate([rng.uniform(-0.2, 2.2, (n - 6, 3)) * [1, 1, 0.2], pts[:6]], axis=0)
    ns = rng.standard_normal((n, 3))
    ns /= np.linalg.norm(ns, axis=1)[:, None]
    n_tris = tris.shape[0]
    pairs = np.stack([rng.integers(0, n, 300), rng.integers(0, n_tris, 300),
                      rng.integers(0, 3, 300)], axis=1)
    return pts, tris, obs, ns, pairs


@pytest.mark.parametrize('K', sorted(KIDS))
@pytest.mark.parametrize('nq,lanes', [(3, 1), (5, 4), (8, 16), (8, 64), (12, 0)])
def test_interior_pairs_match_oracle(oracle, K, nq, lanes):
    from tectosaur_amd import _native as N
    pts, tris, obs, ns, pairs = _case()
    qx, qw = gauss2d_tri(nq)
    params = [3e10, 0.25]
    # device copies held for the whole call (a temporary's memory would be
    # handed to the next allocation)
    d = [_dev(qx), _dev(qw), _dev(obs), _dev(ns), _dev(pts), _dev(tris, torch.int32),
         _dev(pairs, torch.int32)]
    for fp in (0, 1):
        out = torch.empty((pairs.shape[0], 27), dtype=torch.float64, device='cuda')
        print('interior_pairs', K, nq, lanes, fp, flush=True)
        N.call('tct_interior_pairs', KIDS[K], qw.shape[0], N.ptr(d[0]), N.ptr(d[1]),
               N.ptr(d[2]), N.ptr(d[3]), N.ptr(d[4]), N.ptr(d[5]), pairs.shape[0], N.ptr(d[6]),
               0, fp, N.params_array(params).ctypes.data, lanes, N.ptr(out), N.stream_ptr())
        got = out.cpu().numpy().reshape((-1, 3, 3, 3))
        ref = oracle.interior_pairs(K, qx, qw, obs, ns, pts, tris, pairs, False, fp, params)
        scale = np.abs(ref).reshape((ref.shape[0], -1)).max(axis=1)[:, None, None, None]
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale.max())
        assert np.all(np.abs(got - ref) <= 1e-11 * scale + 1e-300)


def test_interior_corners_match_oracle(oracle):
    from tectosaur_amd import _native as N
    pts, tris, obs, ns, _ = _case(1)
    # observation points at a vertex of each triangle they are paired with
    rng = np.random.default_rng(2)
    t = rng.integers(0, tris.shape[0], 200)
    c = rng.integers(0, 3, 200)
    obs_idx = np.arange(200)
    obs_pts = pts[tris[t, c]]
    obs_ns = rng.standard_normal((200, 3))
    pairs = np.stack([obs_idx, t, c], axis=1)
    qt, qw = map_to(gaussxw(10), [0.0, np.pi / 2])
    out = torch.empty((200, 27), dtype=torch.float64, device='cuda')
    d = [_dev(qt), _dev(qw), _dev(obs_ns), _dev(pts), _dev(tris, torch.int32),
         _dev(pairs, torch.int32)]
    print('interior_corners', flush=True)
    N.call('tct_interior_corners', qw.shape[0], N.ptr(d[0]), N.ptr(d[1]), N.ptr(d[2]),
           N.ptr(d[3]), N.ptr(d[4]), 200, N.ptr(d[5]), N.params_array([1.0, 0.25]).ctypes.data,
           N.ptr(out), N.stream_ptr())
    got = out.cpu().numpy().reshape((-1, 3, 3, 3))
    ref = oracle.interior_corners(qt, qw, obs_ns, pts, tris, pairs, [1.0, 0.25])
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize('K', sorted(KIDS))
@pytest.mark.parametrize('float_type', [np.float64, np.float32])
def test_interior_op_matches_oracle(oracle, K, float_type):
    from tectosaur_amd.interior import InteriorOp
    pts, tris, obs, ns, _ = _case(3)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(tris.shape[0] * 9)
    params = [3e10, 0.25]
    print('InteriorOp', K, np.dtype(float_type).name, flush=True)
    op = InteriorOp(obs, ns, (pts, tris), K, 4, 6, 3, 7, params, float_type)
    assert op.vertex_pairs.shape[0] > 0 and op.near_pairs.shape[0] > 0
    got = op.dot(v)
    ref = oracle.interior_op_dot(obs, ns, pts, tris, K, v, params, gauss2d_tri(3),
                                 gauss2d_tri(7), map_to(gaussxw(6), [0.0, np.pi / 2]))
    err = np.linalg.norm(got - ref)
    if float_type == np.float64:
        assert err < 1e-10 * np.linalg.norm(ref), err
    else:
        # the FP32 farfield carries the far-rule terms of the close pairs
        # that the FP64 correction removes: its rounding is relative to them
        close = np.concatenate([op.near_pairs, op.vertex_pairs], axis=0)
        close[:, 2] = 0
        qx, qw = gauss2d_tri(3)
        corr = oracle.pairs_matvec(close, oracle.interior_pairs(K, qx, qw, obs, ns, pts, tris,
                                                                close, False, False, params),
                                   v, obs.shape[0])
        assert err < 2e-5 * (np.linalg.norm(ref) + np.linalg.norm(corr)), err


def test_interior_op_golden():
    # tectosaur tests/test_integral_op.py:200-218 (H, rule 4, FP32)
    from tectosaur_amd.interior import InteriorOp
    corners = [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]]
    pts, tris = mesh.make_rect(3, 3, corners)
    rs = np.random.RandomState(10)
    obs_pts = pts.copy()
    obs_pts[:, 2] += 1.0
    obs_ns = rs.rand(*obs_pts.shape)
    obs_ns /= np.linalg.norm(obs_ns, axis=1)[:, np.newaxis]
    g = np.load(os.path.join(GOLD, 'test_interior.npy'))
    for ft in (np.float32, np.float64):
        op = InteriorOp(obs_pts, obs_ns, (pts, tris), 'elasticH3', 4, 5, 4, 4, [1.0, 0.25], ft)
        out = op.dot(np.ones(tris.shape[0] * 9))
        np.testing.assert_almost_equal(out, g, 6)


def test_interior_rigid_body_near_surface():
    # T of a rigid motion over a closed sphere: -u inside, 0 outside, at
    # points within the near threshold of the surface
    from tectosaur_amd.interior import InteriorOp
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 3)
    rng = np.random.default_rng(5)
    d = rng.standard_normal((400, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = np.where(np.arange(400) % 2 == 0, 0.9, 1.1)
    x = d * r[:, None]
    v = np.tile([0.3, -0.7, 0.5], 3 * tris.shape[0])
    op = InteriorOp(x, d, (pts, tris), 'elasticT3', 4, 8, 3, 8, [1.0, 0.25], np.float64)
    assert op.near_pairs.shape[0] > 400
    y = op.dot(v).reshape((-1, 3))
    expect = np.where((r < 1)[:, None], -np.array([0.3, -0.7, 0.5]), 0.0)
    err = np.abs(y - expect).max()
    print('rigid body max error', err, flush=True)
    assert err < 1e-6
