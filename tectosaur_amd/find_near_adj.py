"""Near-pair finding and adjacency classification (host, one-time setup).

Mirrors tectosaur/mesh/find_near_adj.py (``get_tri_centroids_rs`` :13-20,
``find_close_or_touching`` :29-35) and the classifier
``split_adjacent_close`` of tectosaur/mesh/fast_find_nearfield.cpp:178-217.
The pair search runs in C++ (OpenMP, host_mesh.cpp, reached through the C
ABI), or on the device (near_find.hip: ``find_near_pairs_dev``,
``split_adjacent_close_dev``) with the same pair set and order.
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
    # room for 256 pairs per point (meshes at the usual thresholds have
    # ~90): an overflow runs the search again with the exact count; pages
    # of the buffer that are never written are never committed
    cap = max(16, 256 * ca.shape[0])
    while True:
        out = np.empty((cap, 2), dtype=np.int64)
        N.call('tct_find_near_pairs', ca.shape[0], ca.ctypes.data, ra.ctypes.data, cb.shape[0],
               cb.ctypes.data, rb.ctypes.data, float(threshold), out.ctypes.data, cap,
               ctypes.byref(n))
        if n.value <= cap:
            return out[:n.value].copy()
        cap = n.value


def find_close_or_touching(a_pts, a_tris, b_pts, b_tris, threshold):
    ca, ra = get_tri_centroids_rs(a_pts, a_tris)
    if (a_pts is b_pts or np.array_equal(a_pts, b_pts)) and np.array_equal(a_tris, b_tris):
        cb, rb = ca, ra
    else:
        cb, rb = get_tri_centroids_rs(b_pts, b_tris)
    return find_near_pairs(ca, ra, cb, rb, threshold)


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


def find_near_pairs_dev(ca, ra, cb, rb, threshold, device=None):
    """``find_near_pairs`` on the device: the same pairs in the same order,
    as an int64 device tensor [n, 2].  Inputs are host arrays (the grid is
    sized on the host from the B centroids)."""
    import torch
    ca, ra, cb, rb = (np.ascontiguousarray(x, dtype=np.float64) for x in (ca, ra, cb, rb))
    dev = device if device is not None else N.require_device()
    n_a, n_b = ca.shape[0], cb.shape[0]
    if n_a == 0 or n_b == 0:
        return torch.zeros((0, 2), dtype=torch.int64, device=dev)
    thr = float(threshold)
    rmax_b = float(rb.max())
    lo = np.zeros(3)
    cell = ctypes.c_double(0.0)
    dims = np.zeros(3, dtype=np.int64)
    N.call('tct_near_grid_params', n_b, cb.ctypes.data, thr * (float(ra.max()) + rmax_b) *
           1.0000001, lo.ctypes.data, ctypes.byref(cell), dims.ctypes.data)
    nbytes = ctypes.c_size_t(0)
    N.call('tct_near_pairs_dev_scratch', n_a, n_b, dims.ctypes.data, ctypes.byref(nbytes))
    f64 = dict(dtype=torch.float64, device=dev)
    d_ca, d_ra, d_cb, d_rb = (torch.as_tensor(x, **f64) for x in (ca, ra, cb, rb))
    scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
    offsets = torch.empty(n_a + 1, dtype=torch.int64, device=dev)
    args = (n_a, N.ptr(d_ca), N.ptr(d_ra), n_b, N.ptr(d_cb), N.ptr(d_rb), thr, rmax_b,
            lo.ctypes.data, cell.value, dims.ctypes.data, N.ptr(scratch), nbytes.value)
    st = N.stream_ptr()
    N.call('tct_near_pairs_dev_count', *args, N.ptr(offsets), st)
    total = int(offsets[n_a].item())
    out = torch.empty((total, 2), dtype=torch.int64, device=dev)
    if total:
        N.call('tct_near_pairs_dev_fill', *args, N.ptr(offsets), N.ptr(out), st)
    return out


def split_adjacent_close_dev(close_pairs, tris_A, tris_B):
    """``split_adjacent_close`` on the device: close_pairs an int64 device
    tensor [n, 2], tris int64 device tensors; returns device tensors (near
    [n,2], vertex-adjacent [n,4], edge-adjacent [n,6]) in input order."""
    import torch
    dev = close_pairs.device
    n = close_pairs.shape[0]
    if n == 0:
        return (torch.zeros((0, 2), dtype=torch.int64, device=dev),
                torch.zeros((0, 4), dtype=torch.int64, device=dev),
                torch.zeros((0, 6), dtype=torch.int64, device=dev))
    cp = close_pairs.contiguous()
    ta, tb = tris_A.contiguous(), tris_B.contiguous()
    if cp.dtype != torch.int64 or ta.dtype != torch.int64 or tb.dtype != torch.int64:
        raise TypeError('pairs and triangles must be int64')
    lim = torch.stack([cp[:, 0].max(), cp[:, 1].max(), cp.min()]).cpu().numpy()
    if lim[0] >= ta.shape[0] or lim[1] >= tb.shape[0] or lim[2] < 0:
        raise IndexError('pair index out of range')
    nbytes = ctypes.c_size_t(0)
    N.call('tct_split_adjacent_dev_scratch', n, ctypes.byref(nbytes))
    scratch = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
    near = torch.empty((n, 2), dtype=torch.int64, device=dev)
    va = torch.empty((n, 4), dtype=torch.int64, device=dev)
    ea = torch.empty((n, 6), dtype=torch.int64, device=dev)
    counts = torch.empty(3, dtype=torch.int64, device=dev)
    N.call('tct_split_adjacent_close_dev', n, N.ptr(cp), N.ptr(ta), N.ptr(tb), N.ptr(near),
           N.ptr(va), N.ptr(ea), N.ptr(counts), N.ptr(scratch), nbytes.value, N.stream_ptr())
    c = counts.cpu().numpy()
    return near[:c[0]], va[:c[1]], ea[:c[2]]
