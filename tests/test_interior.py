"""Interior evaluation, CPU: pins of the oracle's point-observer integrals
and the host logic of tectosaur_amd.interior (no GPU needed)."""
import os

import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.quadrature import gauss2d_tri, gaussxw, map_to

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _interior_golden_setup():
    # tectosaur tests/test_integral_op.py:200-218
    corners = [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]]
    pts, tris = mesh.make_rect(3, 3, corners)
    rs = np.random.RandomState(10)
    obs_pts = pts.copy()
    obs_pts[:, 2] += 1.0
    obs_ns = rs.rand(*obs_pts.shape)
    obs_ns /= np.linalg.norm(obs_ns, axis=1)[:, np.newaxis]
    return pts, tris, obs_pts, obs_ns


def _all_pairs(n_obs, n_tris):
    p = np.stack(np.meshgrid(np.arange(n_obs), np.arange(n_tris), indexing='ij'),
                 axis=2).reshape((-1, 2))
    return np.concatenate([p, np.zeros((p.shape[0], 1), np.int64)], axis=1)


def test_oracle_interior_golden(oracle):
    # the fixture is the H operator with the 13-point rule over every
    # triangle, computed in FP32 by the reference
    pts, tris, obs_pts, obs_ns = _interior_golden_setup()
    g = np.load(os.path.join(GOLD, 'test_interior.npy'))
    q4 = gauss2d_tri(4)
    y = oracle.interior_op_dot(obs_pts, obs_ns, pts, tris, 'elasticH3',
                               np.ones(tris.shape[0] * 9), [1.0, 0.25], q4, q4,
                               map_to(gaussxw(5), [0.0, np.pi / 2]))
    np.testing.assert_almost_equal(y, g, 6)
    assert np.max(np.abs(y - g)) < 3e-7
    # a plausible slip: the obs normal taken as the source normal moves it
    e = oracle.interior_pairs('elasticH3', q4[0], q4[1], obs_pts, np.tile([0, 0, 1.0], (9, 1)),
                              pts, tris, _all_pairs(9, tris.shape[0]), True, False, [1.0, 0.25])
    y2 = oracle.pairs_matvec(_all_pairs(9, tris.shape[0]), e, np.ones(tris.shape[0] * 9), 9)
    assert np.max(np.abs(y2 - g)) > 1e-3


@pytest.mark.parametrize('K', ['elasticU3', 'elasticT3', 'elasticA3', 'elasticH3'])
def test_oracle_interior_pairs_equal_point_farfield(oracle, K):
    # two code paths: the pair integrator against the point N-body over the
    # gathered quadrature points (strength J w phi . v)
    rng = np.random.default_rng(5)
    pts, tris = mesh.make_rect(4, 3, [[0, 0, 0], [2, 0, 0.3], [2, 1.5, 0.1], [0, 1.5, 0]])
    n_obs = 7
    obs = rng.uniform(-0.5, 2.5, (n_obs, 3)) + np.array([0, 0, 0.4])
    ns = rng.standard_normal((n_obs, 3))
    ns /= np.linalg.norm(ns, axis=1)[:, None]
    v = rng.standard_normal(tris.shape[0] * 9)
    qx, qw = gauss2d_tri(3)
    allp = _all_pairs(n_obs, tris.shape[0])
    y = oracle.pairs_matvec(allp, oracle.interior_pairs(K, qx, qw, obs, ns, pts, tris, allp,
                                                        True, False, [3e10, 0.25]), v, n_obs)
    phi = np.stack([1 - qx[:, 0] - qx[:, 1], qx[:, 0], qx[:, 1]], axis=1)
    tv = pts[tris]
    un = np.cross(tv[:, 2] - tv[:, 0], tv[:, 2] - tv[:, 1])
    J = np.linalg.norm(un, axis=1)
    sp = np.einsum('qb,tbd->tqd', phi, tv).reshape((-1, 3))
    sn = np.repeat(un / J[:, None], qw.shape[0], axis=0)
    s = np.einsum('t,q,qb,tbd->tqd', J, qw, phi, v.reshape((-1, 3, 3))).reshape((-1, 3))
    y2 = oracle.farfield_pts(K, obs, ns, sp, sn, s, [3e10, 0.25])
    np.testing.assert_allclose(y, y2, rtol=1e-12, atol=1e-12 * np.abs(y2).max())


def test_oracle_rigid_body_identity(oracle):
    # T applied to a rigid motion over any closed surface is -u(x) inside
    # and 0 outside (the Somigliana identity with zero traction)
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    rng = np.random.default_rng(0)
    d = rng.standard_normal((6, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    x = d * np.array([0.3, 0.6, 0.8, 1.3, 1.6, 2.5])[:, None]
    inside = np.linalg.norm(x, axis=1) < 0.9
    qx, qw = gauss2d_tri(8)
    allp = _all_pairs(6, tris.shape[0])
    e = oracle.interior_pairs('elasticT3', qx, qw, x, d, pts, tris, allp, False, False,
                              [1.0, 0.25])
    w = np.array([0.3, -0.5, 0.8])
    motions = [(np.tile(np.eye(3)[c], 3 * tris.shape[0]), np.tile(np.eye(3)[c], (6, 1)))
               for c in range(3)]
    motions.append((np.cross(w[None, :], pts[tris].reshape((-1, 3))).reshape(-1),
                    np.cross(w[None, :], x)))
    for v, u in motions:
        y = oracle.pairs_matvec(allp, e, v, 6).reshape((-1, 3))
        expect = np.where(inside[:, None], -u, 0.0)
        np.testing.assert_allclose(y, expect, atol=1e-7)


def test_oracle_rt_equals_t_on_closed_surface(oracle):
    # integration by parts: summed over a closed surface the surface-curl
    # form RT equals T for a continuous P1 input
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    rng = np.random.default_rng(1)
    v = rng.standard_normal((pts.shape[0], 3))[tris].reshape(-1)
    d = rng.standard_normal((6, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    x = d * np.array([0.4, 0.7, 1.4, 2.0, 0.5, 3.0])[:, None]
    ns = rng.standard_normal((6, 3))
    qx, qw = gauss2d_tri(12)
    allp = _all_pairs(6, tris.shape[0])
    yT = oracle.pairs_matvec(allp, oracle.interior_pairs('elasticT3', qx, qw, x, ns, pts, tris,
                                                         allp, False, False, [1.0, 0.25]), v, 6)
    yR = oracle.pairs_matvec(allp, oracle.interior_pairs('elasticRT3', qx, qw, x, ns, pts, tris,
                                                         allp, False, False, [1.0, 0.25]), v, 6)
    np.testing.assert_allclose(yR, yT, atol=1e-9 * np.abs(yT).max())
    # not an identity for an open surface
    pts2, tris2 = mesh.make_rect(4, 4, [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]])
    v2 = rng.standard_normal((pts2.shape[0], 3))[tris2].reshape(-1)
    allp2 = _all_pairs(6, tris2.shape[0])
    a = oracle.pairs_matvec(allp2, oracle.interior_pairs('elasticT3', qx, qw, x, ns, pts2, tris2,
                                                         allp2, False, False, [1.0, 0.25]), v2, 6)
    b = oracle.pairs_matvec(allp2, oracle.interior_pairs('elasticRT3', qx, qw, x, ns, pts2, tris2,
                                                         allp2, False, False, [1.0, 0.25]), v2, 6)
    assert np.max(np.abs(a - b)) > 1e-3 * np.abs(a).max()


@pytest.mark.parametrize('K', ['elasticH3', 'elasticRT3'])
def test_oracle_rotation_and_finite_part(oracle, K):
    # de-rotation undoes the rotation; finite_part keeps column b = clicks
    rng = np.random.default_rng(2)
    pts, tris = mesh.make_rect(3, 3, [[0, 0, 0], [1, 0, 0], [1, 1, 0.2], [0, 1, 0]])
    x = rng.uniform(0, 1, (3, 3)) + np.array([0, 0, 0.3])
    ns = rng.standard_normal((3, 3))
    qx, qw = gauss2d_tri(5)
    base = np.array([[i, j, 0] for i in range(3) for j in range(tris.shape[0])])
    e0 = oracle.interior_pairs(K, qx, qw, x, ns, pts, tris, base, False, False, [1.0, 0.3])
    for c in (1, 2):
        pr = base.copy()
        pr[:, 2] = c
        ec = oracle.interior_pairs(K, qx, qw, x, ns, pts, tris, pr, False, False, [1.0, 0.3])
        np.testing.assert_allclose(ec, e0, rtol=1e-12, atol=1e-14 * np.abs(e0).max())
        fp = oracle.interior_pairs(K, qx, qw, x, ns, pts, tris, pr, False, True, [1.0, 0.3])
        np.testing.assert_allclose(fp[:, :, c, :], e0[:, :, c, :], rtol=1e-12,
                                   atol=1e-14 * np.abs(e0).max())
        others = [b for b in range(3) if b != c]
        assert np.all(fp[:, :, others, :] == 0.0)


def test_split_vertex_nearfield_matches_oracle(oracle):
    from tectosaur_amd.interior import split_vertex_nearfield
    pts, tris = mesh.make_rect(5, 5, [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]])
    rng = np.random.default_rng(3)
    obs = np.concatenate([pts[::3], rng.uniform(-1, 1, (20, 3)) * [1, 1, 0.1]], axis=0)
    close = oracle.near_pt_pairs(obs, pts, tris, 4.0)
    assert close.shape[0] > 0
    v1, n1 = split_vertex_nearfield(close, obs, pts, tris)
    v2, n2 = oracle.split_vertex_nearfield(close, obs, pts, tris)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(n1, n2)
    assert v1.shape[0] > 0 and n1.shape[0] > 0
    np.testing.assert_allclose(obs[v1[:, 0]], pts[tris[v1[:, 1], v1[:, 2]]], atol=1e-9)


def test_interior_near_pairs_match_oracle(oracle):
    from tectosaur_amd.find_near_adj import find_near_pairs, get_tri_centroids_rs
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    rng = np.random.default_rng(4)
    obs = rng.standard_normal((300, 3)) * 0.7
    got = find_near_pairs(obs, np.zeros(300), *get_tri_centroids_rs(pts, tris), 4.0)
    ref = oracle.near_pt_pairs(obs, pts, tris, 4.0)
    np.testing.assert_array_equal(got, ref)


def test_make_pairs_mat_matches_oracle(oracle):
    from tectosaur_amd.interior import make_pairs_mat
    rng = np.random.default_rng(6)
    pairs = np.array([[0, 1, 0], [2, 0, 0], [0, 3, 1], [1, 1, 2], [2, 0, 0]])
    e = rng.standard_normal((5, 3, 3, 3))
    v = rng.standard_normal(4 * 9)
    m = make_pairs_mat(pairs, e, (3 * 3, 4 * 9))
    np.testing.assert_allclose(m.dot(v), oracle.pairs_matvec(pairs, e, v, 3), rtol=1e-13,
                               atol=1e-13)
