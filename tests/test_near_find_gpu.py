"""Device near-pair finder and classifier (near_find.hip): the same pair
set and order as the host finder, the reference's golden masters, and the
brute-force oracle."""
import os

import numpy as np
import pytest
import torch

from tectosaur_amd import mesh
from tectosaur_amd.find_near_adj import (find_near_pairs, find_near_pairs_dev,
                                         get_tri_centroids_rs, split_adjacent_close,
                                         split_adjacent_close_dev)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden')
SQ = [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]]


def _dev_close(m1, m2, thr):
    return find_near_pairs_dev(*get_tri_centroids_rs(*m1), *get_tri_centroids_rs(*m2), thr)


def _dev_split(close, t1, t2):
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device='cuda')
    return [x.cpu().numpy() for x in split_adjacent_close_dev(close, d(t1), d(t2))]


def test_golden_close_or_touching_dev():
    m = mesh.make_rect(20, 20, SQ)
    out = _dev_close(m, m, 1.25).cpu().numpy()
    g = np.load(os.path.join(GOLD, 'test_close_or_touching.npy'))
    assert np.array_equal(np.sort(out, axis=0), g)


def test_golden_nearfield_real_and_adjacency_dev():
    m = mesh.make_rect(20, 20, SQ)
    close = _dev_close(m, m, 1.25)
    near, va, ea = _dev_split(close, m[1], m[1])
    out = near[np.lexsort([near[:, 1], near[:, 0]])].flatten()
    assert np.array_equal(out, np.load(os.path.join(GOLD, 'test_find_nearfield_real.npy')))
    s = mesh.make_sphere([0, 0, 0], 1.0, 3)
    close = _dev_close(s, s, 1.25)
    near, va, ea = _dev_split(close, s[1], s[1])
    allp = np.zeros((va.shape[0] + ea.shape[0], 4))
    allp[:va.shape[0]] = va
    allp[va.shape[0]:, :2] = ea[:, :2]
    allp = allp[np.lexsort([allp[:, 1], allp[:, 0]])]
    assert np.array_equal(allp, np.load(os.path.join(GOLD, 'test_find_adjacency.npy')))


def test_golden_close_notself_dev():
    m = mesh.make_rect(2, 2, SQ)
    obs = np.array([0.5 * np.ones(5), 0.5 * np.ones(5), np.linspace(0, 2, 5)]).T.copy()
    out = find_near_pairs_dev(obs, np.zeros(5), *get_tri_centroids_rs(*m), 1.0).cpu().numpy()
    g = np.load(os.path.join(GOLD, 'test_find_close_notself.npy'))
    assert np.array_equal(np.sort(out, axis=0), np.sort(g, axis=0))


def _meshes():
    yield 'icosphere4', mesh.make_icosphere([0, 0, 0], 1.0, 4)
    yield 'fault_surface', mesh.fault_surface_mesh(41, 21, 11)[:2]
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1, 1, (900, 3)) * [1, 1, 0.05]
    tris = rng.integers(0, 900, (1500, 3))
    yield 'random_soup', (pts, tris)


@pytest.mark.parametrize('thr', [0.0, 1.0, 2.5, 4.0])
def test_pairs_and_classes_match_host(thr):
    for name, m in _meshes():
        ca, ra = get_tri_centroids_rs(*m)
        host = find_near_pairs(ca, ra, ca, ra, thr)
        dev = find_near_pairs_dev(ca, ra, ca, ra, thr)
        print(name, thr, host.shape[0], flush=True)
        assert np.array_equal(dev.cpu().numpy(), host), name
        h = split_adjacent_close(host, m[1], m[1])
        d = _dev_split(dev, m[1], m[1])
        for a, b in zip(h, d):
            assert np.array_equal(a, b), name


def test_point_observers_and_radii_mix(oracle):
    # zero-radius observers (interior evaluation) and a wide radius spread
    rng = np.random.default_rng(1)
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 3)
    cb, rb = get_tri_centroids_rs(pts, tris)
    rb = rb * np.exp(rng.uniform(-1.5, 1.5, rb.shape[0]))
    ca = rng.standard_normal((3000, 3)) * 0.8
    ra = np.where(rng.random(3000) < 0.5, 0.0, rng.uniform(0, 0.2, 3000))
    for thr in (1.0, 4.0):
        dev = find_near_pairs_dev(ca, ra, cb, rb, thr).cpu().numpy()
        assert np.array_equal(dev, find_near_pairs(ca, ra, cb, rb, thr))
    # brute force on a small case
    dev = find_near_pairs_dev(ca[:300], ra[:300], cb, rb, 2.0).cpu().numpy()
    cap = 1 << 20
    out = np.zeros(2 * cap, dtype=np.int64)
    k = oracle.lib().or_near_pairs_bruteforce(
        300, oracle._p(np.ascontiguousarray(ca[:300])), oracle._p(np.ascontiguousarray(ra[:300])),
        cb.shape[0], oracle._p(np.ascontiguousarray(cb)), oracle._p(np.ascontiguousarray(rb)),
        2.0, oracle._p(out, oracle._i64), cap)
    assert np.array_equal(dev, out[:2 * k].reshape((k, 2)))


def test_single_point_and_empty():
    m = mesh.make_rect(4, 4, SQ)
    cb, rb = get_tri_centroids_rs(*m)
    far = find_near_pairs_dev(np.array([[100.0, 0, 0]]), np.zeros(1), cb, rb, 2.0)
    assert far.shape == (0, 2)
    one = find_near_pairs_dev(cb[:1], rb[:1], cb, rb, 1.0).cpu().numpy()
    assert np.array_equal(one, find_near_pairs(cb[:1], rb[:1], cb, rb, 1.0))
    e = split_adjacent_close_dev(far, torch.zeros((1, 3), dtype=torch.int64, device='cuda'),
                                 torch.as_tensor(m[1].astype(np.int64), device='cuda'))
    assert [x.shape[0] for x in e] == [0, 0, 0]
