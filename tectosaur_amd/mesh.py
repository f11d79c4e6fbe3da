"""Mesh generation for the synthetic workloads (host, numpy, one-time).

Mirrors tectosaur/mesh/mesh_gen.py (``make_rect`` :44-47 with the
topology of :28-41, ``make_sphere`` :59-67), tectosaur/mesh/modify.py
(``concat`` :29-33, ``remove_duplicate_pts`` :7-21 with the index map of
tectosaur/mesh/fast_modify.cpp:15-61) and tectosaur/mesh/refine.py
(``refine`` :13-31).  Triangle and point numbering follow the reference, so
golden fixtures built on these meshes compare index for index.

``make_icosphere`` is new: the nearfield workload asks for an icosphere,
which the reference does not generate.
"""
import numpy as np
import scipy.spatial


def rect_points(corners, xhat, yhat):
    corners = np.asarray(corners, dtype=np.float64)
    X, Y = np.meshgrid(xhat, yhat)
    x = X.reshape(-1)
    y = Y.reshape(-1)
    # bilinear map; corner order as the reference's rect_basis
    w = np.stack([x * y, (1 - x) * y, (1 - x) * (1 - y), x * (1 - y)], axis=1)
    return w @ corners


def rect_topology(nx, ny):
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing='ij')
    i = i.reshape(-1)
    j = j.reshape(-1)
    tl = j * nx + i
    tr = j * nx + i + 1
    bl = (j + 1) * nx + i
    br = (j + 1) * nx + i + 1
    tris = np.empty((2 * i.shape[0], 3), dtype=np.int64)
    tris[0::2] = np.stack([tl, bl, tr], axis=1)
    tris[1::2] = np.stack([bl, br, tr], axis=1)
    return tris


def make_rect(nx, ny, corners):
    """(nx-1)(ny-1) quads split into 2(nx-1)(ny-1) triangles."""
    return rect_points(corners, np.linspace(0, 1, nx), np.linspace(0, 1, ny)), \
        rect_topology(nx, ny)


def remove_duplicate_pts(m, threshold=None):
    pts, tris = m
    if threshold is None:
        rng = np.max(np.max(pts, axis=0) - np.min(pts, axis=0))
        threshold = rng * 1e-13
    pairs = scipy.spatial.cKDTree(pts).query_pairs(threshold, output_type='ndarray')
    n = pts.shape[0]
    dup_of = np.full(n, -1, dtype=np.int64)
    if pairs.shape[0] > 0:
        pairs = np.sort(pairs, axis=1)
        # each dependent (larger index) maps to its smallest earlier duplicate
        order = np.lexsort((pairs[:, 0], pairs[:, 1]))
        pairs = pairs[order]
        first = np.ones(pairs.shape[0], dtype=bool)
        first[1:] = pairs[1:, 1] != pairs[:-1, 1]
        dup_of[pairs[first, 1]] = pairs[first, 0]
    idx_map = np.empty(n, dtype=np.int64)
    cur = 0
    for i in range(n):
        if dup_of[i] >= 0:
            idx_map[i] = idx_map[dup_of[i]]
        else:
            idx_map[i] = cur
            cur += 1
    out_pts = np.empty((cur, pts.shape[1]))
    out_pts[idx_map] = pts  # later duplicates overwrite, as the reference does
    return out_pts, idx_map[tris]


def concat_two_no_remove_duplicates(m1, m2):
    return np.vstack((m1[0], m2[0])), np.vstack((m1[1], m2[1] + m1[0].shape[0]))


def concat(*ms):
    full = ms[0]
    for m in ms[1:]:
        full = remove_duplicate_pts(concat_two_no_remove_duplicates(full, m))
    return full


def refine(m):
    """Split every triangle into 4 at its edge midpoints (refine.py:13-31)."""
    pts, tris = m
    c0, c1, c2 = pts[tris[:, 0]], pts[tris[:, 1]], pts[tris[:, 2]]
    new_pts = np.vstack((pts, (c0 + c1) / 2.0, (c1 + c2) / 2.0, (c2 + c0) / 2.0))
    nt = tris.shape[0]
    f = pts.shape[0]
    i = np.arange(nt)
    m01, m12, m20 = f + i, f + nt + i, f + 2 * nt + i
    new_tris = np.empty((4 * nt, 3), dtype=np.int64)
    new_tris[0::4] = np.stack([tris[:, 0], m01, m20], axis=1)
    new_tris[1::4] = np.stack([tris[:, 1], m12, m01], axis=1)
    new_tris[2::4] = np.stack([tris[:, 2], m20, m12], axis=1)
    new_tris[3::4] = np.stack([m01, m12, m20], axis=1)
    return remove_duplicate_pts((new_pts, new_tris))


def _spherify(pts):
    return pts / np.linalg.norm(pts, axis=1)[:, None]


def make_sphere(center, r, refinements):
    """Octahedron refined ``refinements`` times and projected to the sphere
    (8 * 4**refinements triangles)."""
    pts = np.array([[0, -1.0, 0], [1.0, 0, 0], [0, 0, 1.0], [-1.0, 0, 0], [0, 0, -1.0],
                    [0, 1.0, 0]])
    tris = np.array([[1, 0, 2], [2, 0, 3], [3, 0, 4], [4, 0, 1], [5, 1, 2], [5, 2, 3],
                     [5, 3, 4], [5, 4, 1]])
    m = (pts, tris)
    for _ in range(refinements):
        m = refine(m)
        m = (_spherify(m[0]), m[1])
    return np.array(center) + r * m[0], m[1]


def make_icosphere(center, r, level):
    """Icosahedron refined ``level`` times and projected to the sphere
    (20 * 4**level triangles, outward normals)."""
    t = (1.0 + np.sqrt(5.0)) / 2.0
    pts = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                    [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                   dtype=np.float64)
    tris = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                     [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                     [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                     [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    m = (_spherify(pts), tris)
    for _ in range(level):
        m = refine(m)
        m = (_spherify(m[0]), m[1])
    return np.array(center, dtype=np.float64) + r * m[0], m[1]


def tri_normals(pts, tris):
    """Unscaled normals (v2 - v0) x (v2 - v1) (integral_primitives.cl:24-30)."""
    v = pts[tris]
    return np.cross(v[:, 2] - v[:, 0], v[:, 2] - v[:, 1])


def fault_surface_mesh(n_surf, n_fault_x, n_fault_z, surf_w=20000.0, fault_L=10000.0,
                       fault_W=None, dip_deg=60.0):
    """Flat free surface (n_surf x n_surf vertices over [-surf_w, surf_w]^2,
    z = 0) plus a planar fault whose top edge lies in the surface along the
    x axis and which dips by ``dip_deg`` toward +y.  The fault top edge
    vertices coincide with surface vertices when the spacings match and are
    merged by ``concat``.  Returns (pts, tris, n_surface_tris)."""
    if fault_W is None:
        fault_W = fault_L * (n_fault_z - 1) / (n_fault_x - 1)
    surf = make_rect(n_surf, n_surf, [[-surf_w, -surf_w, 0], [surf_w, -surf_w, 0],
                                      [surf_w, surf_w, 0], [-surf_w, surf_w, 0]])
    dip = np.deg2rad(dip_deg)
    dy = fault_W * np.cos(dip)
    dz = -fault_W * np.sin(dip)
    fault = make_rect(n_fault_x, n_fault_z, [[-fault_L, 0, 0], [fault_L, 0, 0],
                                             [fault_L, dy, dz], [-fault_L, dy, dz]])
    pts, tris = concat(surf, fault)
    return pts, tris, surf[1].shape[0]


def two_surface_mesh():
    """The two-surface farfield workload: a 40 km x 40 km free surface of
    256 x 256 squares (131,072 triangles) plus a 60-degree dipping planar
    fault of 128 x 64 squares (16,384 triangles) whose top edge lies in the
    surface, concatenated with duplicate points merged: 147,456 triangles."""
    pts, tris, n_surf = fault_surface_mesh(257, 129, 65, surf_w=20000.0, fault_L=10000.0,
                                           dip_deg=60.0)
    return pts, tris


def fault_topography_mesh():
    """The large-solve mesh: a 40 km free surface of 512 x 512 squares with
    sinusoidal topography (amplitude 500 m, wavelength 10 km; zero along
    the fault trace) plus a 60-degree dipping fault of 256 x 256 squares
    whose top edge lies in the surface: 655,360 triangles."""
    pts, tris, n_surf = fault_surface_mesh(513, 257, 257, surf_w=20000.0, fault_L=10000.0,
                                           dip_deg=60.0)
    surf = np.zeros(pts.shape[0], dtype=bool)
    surf[np.unique(tris[:n_surf])] = True
    lam = 10000.0
    pts[surf, 2] += 500.0 * np.sin(2 * np.pi * pts[surf, 0] / lam) * \
        np.sin(2 * np.pi * pts[surf, 1] / lam)
    return pts, tris
