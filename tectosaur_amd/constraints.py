"""Linear constraints on the DOF vector (host, one-time setup per mesh).

Mirrors tectosaur/constraints.py (``build_constraint_matrix`` :10-17, the
``Term`` / ``ConstraintEQ`` records of tectosaur/fast_constraints.cpp:13-40),
tectosaur/continuity.py (``find_touching_pts`` :7-12, ``tri_side`` :37-50,
``get_side_of_fault`` :52-74, ``continuity_constraints`` :79-143) and
tectosaur/constraint_builders.py (``find_free_edges`` :74-91,
``free_edge_dofs`` :96-107, ``free_edge_constraints`` :120-138).  The
elimination itself runs in C++ (host_constraints.cpp) through the C ABI.

The GMRES operator of a constrained solve is cm^T A cm, and the solution is
u = cm x + c_rhs (examples/solve.py:25-61 of the reference).
"""
import collections
import ctypes

import numpy as np
import scipy.sparse

from . import _native as N

Term = collections.namedtuple('Term', ['val', 'dof'])
ConstraintEQ = collections.namedtuple('ConstraintEQ', ['terms', 'rhs'])


def build_constraint_matrix(cs, n_total_dofs):
    """(cm [n_total, n_total - n_eliminated] CSR, c_rhs [n_total], rhs_mat)."""
    n_cs = len(cs)
    n_terms = np.array([len(c.terms) for c in cs], dtype=np.int64)
    dofs = np.array([t.dof for c in cs for t in c.terms], dtype=np.int64)
    vals = np.array([t.val for c in cs for t in c.terms], dtype=np.float64)
    rhs_in = np.array([c.rhs for c in cs], dtype=np.float64)
    cm_n, r_n, n_el = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    cap_cm = n_total_dofs + 4 * max(len(dofs), 1)
    cap_r = 4 * max(n_cs, 1)
    for _ in range(2):
        cmr = np.empty(cap_cm, np.int64); cmc = np.empty(cap_cm, np.int64)
        cmv = np.empty(cap_cm, np.float64)
        rr = np.empty(cap_r, np.int64); rc = np.empty(cap_r, np.int64)
        rv = np.empty(cap_r, np.float64)
        N.call('tct_build_constraint_matrix', n_cs, n_terms.ctypes.data, dofs.ctypes.data,
               vals.ctypes.data, n_total_dofs, cap_cm, cmr.ctypes.data, cmc.ctypes.data,
               cmv.ctypes.data, ctypes.byref(cm_n), cap_r, rr.ctypes.data, rc.ctypes.data,
               rv.ctypes.data, ctypes.byref(r_n), ctypes.byref(n_el))
        if cm_n.value <= cap_cm and r_n.value <= cap_r:
            break
        cap_cm, cap_r = cm_n.value, max(r_n.value, 1)
    k, kr = cm_n.value, r_n.value
    n_cols = n_total_dofs - n_el.value
    cm = scipy.sparse.csr_matrix((cmv[:k], (cmr[:k], cmc[:k])), shape=(n_total_dofs, n_cols))
    rhs_mat = scipy.sparse.csr_matrix((rv[:kr], (rr[:kr], rc[:kr])), shape=(n_total_dofs, n_cs))
    return cm, rhs_mat.dot(rhs_in) if n_cs else np.zeros(n_total_dofs), rhs_mat


def find_touching_pts(tris):
    out = [[] for _ in range(int(np.max(tris)) + 1)]
    for i, t in enumerate(tris):
        for d in range(3):
            out[t[d]].append((i, d))
    return out


def tri_side(tri1, tri2, threshold=1e-12):
    n = np.cross(tri1[2] - tri1[0], tri1[2] - tri1[1])
    n = n / np.linalg.norm(n)
    direction = np.mean(tri2, axis=0) - np.mean(tri1, axis=0)
    direction = direction / np.linalg.norm(direction)
    dot_val = direction.dot(n)
    if dot_val > threshold:
        return 0
    elif dot_val < -threshold:
        return 1
    return 2


def get_side_of_fault(pts, tris, fault_start_idx):
    """Side (1 or 2, 0 = not touching the fault) of every surface triangle
    that shares a vertex with a fault triangle (continuity.py:52-74)."""
    n = tris.shape[0]
    touching = [[] for _ in range(int(np.max(tris)) + 1)]
    for i in range(n):
        for d in range(3):
            touching[tris[i, d]].append(i)
    side = np.zeros(n)
    shared_verts = np.zeros(n)
    for lst in touching:
        for row in lst:
            if row >= fault_start_idx:
                continue
            for col in lst:
                if col < fault_start_idx:
                    continue
                surf_tri, fault_tri = tris[row], tris[col]
                which = tri_side(pts[fault_tri], pts[surf_tri])
                n_shared = sum(1 for d in range(3) if surf_tri[d] in fault_tri)
                if shared_verts[row] < 2:
                    side[row] = int(which) + 1
                    shared_verts[row] = n_shared
    return side


def continuity_constraints(pts, tris, fault_start_idx, tensor_dim=3):
    """Displacement continuity between triangles sharing a vertex of the
    surface part tris[:fault_start_idx]; across the fault the jump equals the
    fault slip (continuity.py:79-143)."""
    surface_tris = tris[:fault_start_idx]
    fault_tris = tris[fault_start_idx:]
    touching_pt = find_touching_pts(surface_tris)
    side = get_side_of_fault(pts, tris, fault_start_idx)
    cs = []
    for tpt in touching_pt:
        for ii in range(len(tpt)):
            it_idx, ic_idx = tpt[ii]
            for jj in range(ii + 1, len(tpt)):
                dt_idx, dc_idx = tpt[jj]
                dependent_tri = surface_tris[dt_idx]
                side1, side2 = side[it_idx], side[dt_idx]
                crosses = (side1 != side2) and (side1 != 0) and (side2 != 0)
                fault_tri_idx = None
                if crosses:
                    fti, fci = np.where(fault_tris == dependent_tri[dc_idx])
                    if fti.shape[0] != 0:
                        fault_tri_idx, fault_corner_idx = fti[0], fci[0]
                for d in range(tensor_dim):
                    ind = (it_idx * 3 + ic_idx) * tensor_dim + d
                    dep = (dt_idx * 3 + dc_idx) * tensor_dim + d
                    if dep <= ind:
                        continue
                    terms = [Term(1.0, dep), Term(-1.0, ind)]
                    if fault_tri_idx is not None:
                        fdof = fault_start_idx * 9 + fault_tri_idx * 9 + fault_corner_idx * 3 + d
                        terms.append(Term(-1.0 if side1 < side2 else 1.0, fdof))
                    cs.append(ConstraintEQ(terms, 0.0))
    return cs


def find_free_edges(tris):
    """(triangle, local edge) of every edge used by exactly one triangle, in
    the order the reference's dict walk yields them (first appearance)."""
    t = np.asarray(tris, dtype=np.int64)
    if t.shape[0] == 0:
        return []
    a = t.reshape(-1)
    b = t[:, [1, 2, 0]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * (int(t.max()) + 1) + hi
    _, first, cnt = np.unique(key, return_index=True, return_counts=True)
    idx = np.sort(first[cnt == 1])
    return [(int(i // 3), int(i % 3)) for i in idx]


def free_edge_dofs(tris, free_edges):
    pt_idxs = set()
    for tri_idx, edge_idx in free_edges:
        for v in range(2):
            pt_idxs.add(int(tris[tri_idx][(edge_idx + v) % 3]))
    return [i * 3 + v for i, t in enumerate(tris) for v in range(3) if int(t[v]) in pt_idxs]


def free_edge_constraints(tris, field=None):
    """Pin every DOF at a vertex of a free edge to ``field`` (default 0)."""
    if field is None:
        field = np.zeros(tris.shape[0] * 9)
    cs = []
    for dof in free_edge_dofs(tris, find_free_edges(tris)):
        for d in range(3):
            vd = dof * 3 + d
            cs.append(ConstraintEQ([Term(1.0, vd)], field[vd]))
    return cs
