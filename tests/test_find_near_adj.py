"""Near-pair finder (host C++ through the C ABI) and adjacency classifier,
pinned by the reference's golden masters and exact assertions
(tectosaur tests/test_find_near_adj.py) and by the brute-force oracle."""
import os

import numpy as np
import pytest

from tectosaur_amd import mesh
from tectosaur_amd.find_near_adj import (find_close_or_touching, split_adjacent_close,
                                         get_tri_centroids_rs, find_near_pairs)

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
SQ = [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]]


def test_find_adjacents():
    # tests/test_find_near_adj.py:12-21 (the va assertion; ea now carries 6 columns)
    tris = np.array([[0, 1, 2], [2, 1, 3], [0, 4, 5]])
    pts = np.random.default_rng(0).random((6, 3))
    close = find_close_or_touching(pts, tris, pts, tris, 1.0)
    near, va, ea = split_adjacent_close(close, tris, tris)
    assert va.size == 8
    assert np.all(va.flatten() == (0, 2, 0, 0, 2, 0, 0, 0))
    assert np.all(ea[:, :2] == [[0, 1], [1, 0]])
    assert np.all(ea[0, 2:] == (1, 1, 2, 0)) and np.all(ea[1, 2:] == (0, 2, 1, 1))


def test_nearfield_exact():
    # tests/test_find_near_adj.py:23-37
    pts, tris = mesh.make_rect(3, 3, SQ)
    assert tris.shape[0] == 8
    close = find_close_or_touching(pts, tris, pts, tris, 1.0)
    near, va, ea = split_adjacent_close(close, tris, tris)
    check_for = [(0, 5), (0, 6), (0, 3), (1, 7), (2, 7), (3, 0), (4, 7), (5, 0), (6, 0),
                 (7, 2), (7, 1), (7, 4)]
    got = [tuple(p) for p in near]
    assert len(got) == len(check_for)
    for p in check_for:
        assert p in got


def test_close_or_touching_golden():
    pts, tris = mesh.make_rect(20, 20, SQ)
    out = np.sort(find_close_or_touching(pts, tris, pts, tris, 1.25), axis=0)
    g = np.load(os.path.join(GOLD, 'test_close_or_touching.npy'))
    assert np.array_equal(out, g)


def test_find_nearfield_real_golden():
    pts, tris = mesh.make_rect(20, 20, SQ)
    close = find_close_or_touching(pts, tris, pts, tris, 1.25)
    near, va, ea = split_adjacent_close(close, tris, tris)
    out = near[np.lexsort([near[:, 1], near[:, 0]])].flatten()
    assert np.array_equal(out, np.load(os.path.join(GOLD, 'test_find_nearfield_real.npy')))


def test_find_adjacency_golden():
    # the fixture stores vertex-adjacent rows (i, j, ci, cj) and edge-adjacent pairs (i, j, 0, 0)
    m = mesh.make_sphere([0, 0, 0], 1.0, 3)
    close = find_close_or_touching(*m, *m, 1.25)
    near, va, ea = split_adjacent_close(close, m[1], m[1])
    allp = np.zeros((va.shape[0] + ea.shape[0], 4))
    allp[:va.shape[0]] = va
    allp[va.shape[0]:, :2] = ea[:, :2]
    allp = allp[np.lexsort([allp[:, 1], allp[:, 0]])]
    g = np.load(os.path.join(GOLD, 'test_find_adjacency.npy'))
    assert np.array_equal(allp, g)


def test_find_close_notself_golden():
    m = mesh.make_rect(2, 2, SQ)
    obs = np.array([0.5 * np.ones(5), 0.5 * np.ones(5), np.linspace(0, 2, 5)]).T.copy()
    out = find_near_pairs(obs, np.zeros(5), *get_tri_centroids_rs(*m), 1.0)
    g = np.load(os.path.join(GOLD, 'test_find_close_notself.npy'))
    assert np.array_equal(np.sort(out, axis=0), np.sort(g, axis=0))


@pytest.mark.parametrize('thr', [0.5, 1.25, 3.0])
def test_matches_bruteforce_oracle(oracle, thr):
    rng = np.random.default_rng(int(thr * 10))
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    pts = pts + 0.01 * rng.normal(size=pts.shape)
    got = find_close_or_touching(pts, tris, pts, tris, thr)
    want = oracle.near_pairs(pts, tris, pts, tris, thr)
    assert np.array_equal(got, want)
    # and different obs / src sets
    got = find_close_or_touching(pts, tris[:100], pts, tris[50:], thr)
    want = oracle.near_pairs(pts, tris[:100], pts, tris[50:], thr)
    assert np.array_equal(got, want)


def test_classifier_matches_oracle(oracle):
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    close = find_close_or_touching(pts, tris, pts, tris, 2.0)
    for a, b in zip(split_adjacent_close(close, tris, tris),
                    oracle.split_adjacent_close(close, tris, tris)):
        assert np.array_equal(a, b)


def test_empty_inputs():
    out = find_near_pairs(np.zeros((0, 3)), np.zeros(0), np.zeros((3, 3)), np.ones(3), 1.0)
    assert out.shape == (0, 2)
    near, va, ea = split_adjacent_close(np.zeros((0, 2), np.int64), np.zeros((0, 3)),
                                        np.zeros((0, 3)))
    assert near.shape == (0, 2) and va.shape == (0, 4) and ea.shape == (0, 6)


def test_resolve_ea_rotation_matches_oracle(oracle):
    from tectosaur_amd.nearfield_op import resolve_ea_rotation
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    tris2 = tris.copy()
    tris2[::3] = tris2[::3][:, [1, 0, 2]]  # inconsistent orientation -> flips
    for t in (tris, tris2):
        close = find_close_or_touching(pts, t, pts, t, 1.5)
        _, _, ea = split_adjacent_close(close, t, t)
        assert np.array_equal(resolve_ea_rotation(t, ea), oracle.resolve_ea_rotation(t, ea))
