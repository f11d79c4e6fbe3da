"""Near-pair finding and adjacency classification (host, one-time setup).

Mirrors tectosaur/mesh/find_near_adj.py (``get_tri_centroids_rs`` :13-20,
``find_close_or_touching`` :29-35) and the classifier
``split_adjacent_close`` of tectosaur/mesh/fast_find_nearfield.cpp:178-217.
The pair search runs in C++ (OpenMP, host_mesh.cpp, reached through the C
ABI); the classifier is vectorised numpy.
"""
import ctypes

import numpy as np

from . import _native as N


def get_tri_centroids_rs(pts, tris):
    tri_pts = pts[tris]
    centroid = np.sum(tri_pts, axis=1) / 3.0
    r = np.sqrt(np.max(np.sum((tri_pts - centroid[:, np.newaxis, :]) ** 2, axis=2), axis=1))
    return centroid, r


def find_near_pairs(ca, ra, cb, rb, threshold):
    """All (i, j), i into a and j into b, with
    |c_i - c_j|^2 <= ((r_i + r_j) * threshold)^2, sorted by (i, j)."""
    ca, ra, cb, rb = (np.ascontiguousarray(x, dtype=np.float64) for x in (ca, ra, cb, rb))
    n = ctypes.c_int64(0)
    cap = max(16, 64 * ca.shape[0])
    while True:
        out = np.empty((cap, 2), dtype=np.int64)
        N.call('tct_find_near_pairs', ca.shape[0], ca.ctypes.data, ra.ctypes.data, cb.shape[0],
               cb.ctypes.data, rb.ctypes.data, float(threshold), out.ctypes.data, cap,
               ctypes.byref(n))
        if n.value <= cap:
            return out[:n.value].copy()
        cap = n.value


def find_close_or_touching(a_pts, a_tris, b_pts, b_tris, threshold):
    return find_near_pairs(*get_tri_centroids_rs(a_pts, a_tris),
                           *get_tri_centroids_rs(b_pts, b_tris), threshold)


def split_adjacent_close(close_pairs, tris_A, tris_B):
    """(near [n,2], vertex-adjacent [n,4] (i, j, corner_i, corner_j),
    edge-adjacent [n,6] (i, j, corner pair 1, corner pair 2)); pairs sharing
    all three vertices (coincident) are dropped.  Shared corners are listed
    in (corner_i, corner_j) row-major order, as the reference's loop does.
    Runs in C++ (host_mesh.cpp) through the C ABI."""
    cp = np.ascontiguousarray(np.asarray(close_pairs, dtype=np.int64).reshape((-1, 2)))
    n = cp.shape[0]
    if n == 0:
        return (np.zeros((0, 2), np.int64), np.zeros((0, 4), np.int64),
                np.zeros((0, 6), np.int64))
    ta = np.ascontiguousarray(np.asarray(tris_A, dtype=np.int64))
    tb = np.ascontiguousarray(np.asarray(tris_B, dtype=np.int64))
    if cp[:, 0].max() >= ta.shape[0] or cp[:, 1].max() >= tb.shape[0] or cp.min() < 0:
        raise IndexError('pair index out of range')
    near = np.empty((n, 2), np.int64)
    va = np.empty((n, 4), np.int64)
    ea = np.empty((n, 6), np.int64)
    counts = np.zeros(3, np.int64)
    N.call('tct_split_adjacent_close', n, cp.ctypes.data, ta.ctypes.data, tb.ctypes.data,
           near.ctypes.data, va.ctypes.data, ea.ctypes.data, counts.ctypes.data)
    return near[:counts[0]].copy(), va[:counts[1]].copy(), ea[:counts[2]].copy()
