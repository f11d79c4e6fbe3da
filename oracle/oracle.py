"""CPU oracle for the farfield / nearfield hot path -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module, and only as the checker.  The product
package ``tectosaur_amd`` never imports it.

It loads ``liboracle.so`` (built from tct_oracle.c by the Makefile here) and
adds numpy restatements of the host-side reference steps.

Pinning (what fixes each function independently of this restatement):

  kernel_tensor / kernel_vector   U, T, A, H: sympy derivation of the Kelvin
      solution and its traction operators (tests/golden/make_kernel_golden.py,
      the method of tectosaur tests/elastic_derive.py:47-147), plus the
      reference's own tensor==vector identity.
  farfield_pts                    brute-force sums of the pinned point
      kernels; scaling law K(lambda x) = lambda^(scale_type) K(x).
  pairs_integrate (U/T/A/H)       golden master tests/golden_masters/
      test_farfield_two_tris.npy (H, nq=3); test_gpu_vert_adjacent.npy
      (H, vertex-adjacent, nq=3); refinement additivity (a parent coincident
      integral equals the sum of its 16 child-pair integrals).
  pairs_integrate (edge-adjacent, flipped source)  the six numberings of
      one source triangle give the same block per source vertex up to the
      normal's sign; the reference's de-rotation of a flipped source
      (assemble.cl:85-89) departs from that for nonzero clicks, and the
      oracle reproduces exactly the computed departure
      (tests/test_oracle.py::test_edge_adjacent_numbering_invariance).
  pairs_integrate (RT/RA/RH)      no reference fixture exists; pinned by
      the mathematics instead: on two disjoint closed icospheres integration
      by parts makes RT = T, cm^T RA = cm^T A and cm^T RH = cm^T H (to
      <= 2e-9 at order 6, spectrally falling with order), so a wrong sign,
      constant or curl index fails; plus refinement additivity and
      convergence in n.
  farfield_tris                   equals the dense pair integral of the same
      rule summed against the input (two independent code paths), plus the
      golden two-triangle fixture through that identity.
  near_pairs / split_adjacent     golden masters test_close_or_touching.npy,
      test_find_nearfield_real.npy, test_find_adjacency.npy and the exact
      assertions of tectosaur tests/test_find_near_adj.py:18-37.
  bcoo_matvec                     dense matrix-vector product
      (tectosaur tests/test_sparse.py:8-73 semantics).
  interior_pairs (U/T/A/H)        golden master tests/golden_masters/
      test_interior.npy (H, nq=4 over a 2x2 plate, obs normals from the
      seeded RNG of tectosaur tests/test_integral_op.py:200-218) to 2e-7;
      equals farfield_pts over the gathered quadrature points (a separate
      code path); T on a closed sphere reproduces the rigid-body identity
      (-u inside, 0 outside).
  interior_pairs (RT)             the integration-by-parts identity: summed
      over a closed surface, RT equals T for a continuous P1 input.
  interior_pairs_sinh             equals the converged sum over a 4^6-way
      subdivision of the triangle (brute force) for points above, beside
      and near the edges of a triangle; the rigid-body identity at points
      1e-3 triangle sizes from a closed sphere.
  corner_radial / interior_corners  the radial integrals I1, I2, I3 equal an
      independent sympy derivation of the eps-regularised finite parts
      (tests/golden/make_corner_golden.py, cross-checked there by mpmath
      quadrature) to 1e-12 at 24 random (a, rho_max, n); I4 is zero as the
      reference's integer division makes it (the derived I4 is not).  The
      whole corner block equals an adaptive-quadrature composition of those
      derived integrals with the P1 basis, for each vertex of the source
      triangle (tests/test_oracle.py::test_interior_corner_blocks).
  interior_op_dot                 composed from the pinned pieces above;
      the test_interior.npy fixture pins it whole when nq_far == nq_near.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KERNEL_IDS = {
    'elasticU3': 0, 'elasticT3': 1, 'elasticA3': 2, 'elasticH3': 3,
    'elasticRT3': 4, 'elasticRA3': 5, 'elasticRH3': 6,
}

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: run make -C oracle')
        L = ctypes.CDLL(path)
        L.or_kernel_tensor.argtypes = [ctypes.c_int, _d, _d, _d, _d, _d]
        L.or_kernel_vector.argtypes = [ctypes.c_int, _d, _d, _d, _d, _d, _d]
        L.or_farfield_pts.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, ctypes.c_int64,
                                      _d, _d, _d, _d, _i64, ctypes.c_int64, _d]
        L.or_pairs_integrate.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _d, _i32,
                                         ctypes.c_int64, _i64, ctypes.c_int, ctypes.c_int,
                                         _d, _d]
        L.or_farfield_tris.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _d,
                                       ctypes.c_int64, _i32, ctypes.c_int64, _i32, _d, _d,
                                       _i64, ctypes.c_int64, _d]
        L.or_near_pairs_bruteforce.argtypes = [ctypes.c_int64, _d, _d, ctypes.c_int64, _d, _d,
                                               ctypes.c_double, _i64, ctypes.c_int64]
        L.or_near_pairs_bruteforce.restype = ctypes.c_int64
        L.or_bcoo_matvec.argtypes = [ctypes.c_int64, ctypes.c_int, _i64, _i64, _d, _d, _d]
        L.or_interior_pairs.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, _d, _d, _d, _i32,
                                        ctypes.c_int64, _i64, ctypes.c_int, ctypes.c_int, _d, _d]
        L.or_interior_pairs_sinh.argtypes = [ctypes.c_int, ctypes.c_int64, _d, _d, ctypes.c_int64,
                                             _d, _d, _d, _d, _d, _i32, ctypes.c_int64, _i64, _d,
                                             _d]
        L.or_interior_corners.argtypes = [ctypes.c_int64, _d, _d, _d, _d, _i32, ctypes.c_int64,
                                          _i64, _d, _d]
        L.or_corner_radial.argtypes = [_d, ctypes.c_double, _d, _d]
        _LIB = L
    return _LIB


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def kernel_tensor(K, D, m, n, params):
    """3x3 kernel K[d_obs, d_src] at displacement D = src - obs."""
    out = np.zeros(9)
    D, m, n, params = (_f64(a) for a in (D, m, n, params))
    lib().or_kernel_tensor(KERNEL_IDS[K], _p(D), _p(m), _p(n), _p(params), _p(out))
    return out.reshape(3, 3)


def kernel_vector(K, D, m, n, vin, params):
    out = np.zeros(3)
    D, m, n, vin, params = (_f64(a) for a in (D, m, n, vin, params))
    lib().or_kernel_vector(KERNEL_IDS[K], _p(D), _p(m), _p(n), _p(vin), _p(params), _p(out))
    return out


def farfield_pts(K, obs_pts, obs_ns, src_pts, src_ns, vec, params, rows=None):
    """a1: out[i] = sum_j K(y_j - x_i) in_j (r2 == 0 skipped)."""
    obs_pts, obs_ns, src_pts, src_ns, vec, params = (
        _f64(a) for a in (obs_pts, obs_ns, src_pts, src_ns, vec, params))
    n_obs, n_src = obs_pts.shape[0], src_pts.shape[0]
    if rows is None:
        out = np.zeros(n_obs * 3)
        rp, nr = None, 0
    else:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros(rows.shape[0] * 3)
        rp, nr = _p(rows, _i64), rows.shape[0]
    lib().or_farfield_pts(KERNEL_IDS[K], n_obs, _p(obs_pts), _p(obs_ns), n_src, _p(src_pts),
                          _p(src_ns), _p(vec), _p(params), rp, nr, _p(out))
    return out


def pairs_integrate(K, qx, qw, pts, tris, pairs, check0, params):
    """a5: [n, 3, 3, 3, 3] pair integrals; pairs has 2 or 5 columns."""
    qx, qw, pts, params = (_f64(a) for a in (qx, qw, pts, params))
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    pairs = np.ascontiguousarray(pairs, dtype=np.int64)
    n = pairs.shape[0]
    out = np.zeros(n * 81)
    if n > 0:
        lib().or_pairs_integrate(KERNEL_IDS[K], qx.shape[0], _p(qx), _p(qw), _p(pts),
                                 _p(tris, _i32), n, _p(pairs, _i64), pairs.shape[1],
                                 int(bool(check0)), _p(params), _p(out))
    return out.reshape((n, 3, 3, 3, 3))


def farfield_tris(K, qx2, qw2, pts, obs_tris, src_tris, vin, params, rows=None):
    """a2: matrix-free tri-to-tri farfield, out[n_obs * 9]."""
    qx2, qw2, pts, vin, params = (_f64(a) for a in (qx2, qw2, pts, vin, params))
    obs_tris = np.ascontiguousarray(obs_tris, dtype=np.int32)
    src_tris = np.ascontiguousarray(src_tris, dtype=np.int32)
    n_obs = obs_tris.shape[0]
    if rows is None:
        out = np.zeros(n_obs * 9)
        rp, nr = None, 0
    else:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros(rows.shape[0] * 9)
        rp, nr = _p(rows, _i64), rows.shape[0]
    lib().or_farfield_tris(KERNEL_IDS[K], qx2.shape[0], _p(qx2), _p(qw2), _p(pts), n_obs,
                           _p(obs_tris, _i32), src_tris.shape[0], _p(src_tris, _i32),
                           _p(vin), _p(params), rp, nr, _p(out))
    return out


def tri_centroids_rs(pts, tris):
    """Centroid and bounding radius per triangle, the arithmetic of
    tectosaur/mesh/find_near_adj.py:13-20."""
    tri_pts = pts[tris]
    centroid = np.sum(tri_pts, axis=1) / 3.0
    r = np.sqrt(np.max(np.sum((tri_pts - centroid[:, np.newaxis, :]) ** 2, axis=2), axis=1))
    return centroid, r


def near_pairs(a_pts, a_tris, b_pts, b_tris, threshold):
    """All (i, j) with |c_i - c_j|^2 <= ((r_i + r_j) thr)^2, brute force,
    row-major order."""
    ca, ra = tri_centroids_rs(a_pts, a_tris)
    cb, rb = tri_centroids_rs(b_pts, b_tris)
    ca, ra, cb, rb = (_f64(x) for x in (ca, ra, cb, rb))
    cap = 1 << 16
    while True:
        out = np.zeros(2 * cap, dtype=np.int64)
        k = lib().or_near_pairs_bruteforce(ca.shape[0], _p(ca), _p(ra), cb.shape[0], _p(cb),
                                           _p(rb), float(threshold), _p(out, _i64), cap)
        if k <= cap:
            return out[:2 * k].reshape((k, 2))
        cap = k


def split_adjacent_close(close_pairs, tris_a, tris_b):
    """Classify pairs by shared vertices, tectosaur/mesh/fast_find_nearfield.cpp
    :178-217: returns (near [n,2], vertex-adj [n,4], edge-adj [n,6]);
    coincident pairs (3 shared) are dropped."""
    near, va, ea = [], [], []
    for i1, i2 in close_pairs:
        p1 = p2 = None
        co = False
        for d1 in range(3):
            for d2 in range(3):
                if tris_a[i1, d1] != tris_b[i2, d2]:
                    continue
                if p1 is None:
                    p1 = (d1, d2)
                elif p2 is None:
                    p2 = (d1, d2)
                else:
                    co = True
        if co:
            continue
        if p1 is None:
            near.append((i1, i2))
        elif p2 is None:
            va.append((i1, i2, p1[0], p1[1]))
        else:
            ea.append((i1, i2, p1[0], p1[1], p2[0], p2[1]))
    return (np.array(near, dtype=np.int64).reshape((-1, 2)),
            np.array(va, dtype=np.int64).reshape((-1, 4)),
            np.array(ea, dtype=np.int64).reshape((-1, 6)))


def edge_adj_orient(touching_verts):
    """tectosaur/nearfield/nearfield_op.py:44-50."""
    tv = sorted(touching_verts)
    if tv[0] == 0:
        return 2 if tv[1] == 2 else 0
    return 1


def resolve_ea_rotation(tris, ea):
    """tectosaur/nearfield/nearfield_op.py:52-62: (obs, src, obs_clicks,
    src_clicks, src_flip) per edge-adjacent pair."""
    out = []
    for row in ea:
        oc = edge_adj_orient([row[2], row[4]])
        sc = edge_adj_orient([row[3], row[5]])
        flip = (tris[row[0], (0 + oc) % 3] != tris[row[1], (1 + sc) % 3] or
                tris[row[0], (1 + oc) % 3] != tris[row[1], (0 + sc) % 3])
        out.append((row[0], row[1], oc, sc, int(flip)))
    return np.array(out, dtype=np.int64).reshape((-1, 5))


def bcoo_matvec(rows, cols, data, x, shape):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    data = _f64(data)
    x = _f64(x)
    bs = data.shape[1]
    y = np.zeros(shape[0])
    lib().or_bcoo_matvec(rows.shape[0], bs, _p(rows, _i64), _p(cols, _i64), _p(data), _p(x),
                         _p(y))
    return y


def nearfield_blocks(pts, tris, nq_co, nq_ea, nq_va, nq_far, nq_near, near_threshold,
                     K_near, K_far, params, rules, obs_subset=None, src_subset=None):
    """The nearfield assembly of tectosaur/nearfield/nearfield_op.py:95-141,
    computed with the oracle pair integrator, for any obs/src triangle
    subsets (default: all triangles).  ``rules`` supplies the quadrature
    rules (a dict with 'co', 'ea', 'va', 'near', 'far' -> (qx, qw)) so this
    module does not depend on the product package.  Returns a dict class ->
    (pairs in triangle space, pairs in DOF space (positions within the
    subsets), corrected blocks [n,9,9], uncorrected blocks [n,9,9])."""
    n = tris.shape[0]
    obs_subset = np.arange(n) if obs_subset is None else np.asarray(obs_subset, dtype=np.int64)
    src_subset = np.arange(n) if src_subset is None else np.asarray(src_subset, dtype=np.int64)

    def dof_of(tri_pairs):
        # nearfield_op.py:24-38: positions by search, one pair at a time
        out = [[np.where(obs_subset == a)[0][0], np.where(src_subset == b)[0][0]]
               for a, b in tri_pairs]
        return np.array(out, dtype=np.int64).reshape((-1, 2))

    def tri_of(dof_pairs):
        # nearfield_op.py:40-42
        dp = np.asarray(dof_pairs, dtype=np.int64)
        t = np.stack([obs_subset[dp[:, 0]], src_subset[dp[:, 1]]], axis=1)
        return np.concatenate([t, dp[:, 2:]], axis=1)

    co_tris = np.intersect1d(obs_subset, src_subset)
    co = np.stack([co_tris, co_tris], axis=1)
    co_dofs = dof_of(co)
    close = near_pairs(pts, tris[obs_subset], pts, tris[src_subset], near_threshold)
    near_d, va_d, ea_d = split_adjacent_close(close, tris[obs_subset], tris[src_subset])
    near = tri_of(near_d)
    va = tri_of(va_d)
    va = np.concatenate([va, np.zeros((va.shape[0], 1), dtype=np.int64)], axis=1)
    ea = resolve_ea_rotation(tris, tri_of(ea_d))
    out = {}
    for name, prs, dofs, q, check0 in (
            ('co', co, co_dofs, rules['co'], True),
            ('ea', ea, ea_d[:, :2], rules['ea'], True),
            ('va', va, va_d[:, :2], rules['va'], True),
            ('near', near, near_d[:, :2], rules['near'], False)):
        acc = pairs_integrate(K_near, q[0], q[1], pts, tris, prs, check0, params)
        corr = pairs_integrate(K_far, rules['far'][0], rules['far'][1], pts, tris,
                               prs[:, :2], name == 'co', params)
        out[name] = (prs, dofs, (acc - corr).reshape((-1, 9, 9)), acc.reshape((-1, 9, 9)))
    return out


def nearfield_dense(blocks, n_obs, n_src, corrected=True):
    """Dense [n_obs*9, n_src*9] matrix of ``nearfield_blocks``' output
    (the sum of the reference's four BCOO matrices, nearfield_op.py:135-150)."""
    M = np.zeros((n_obs * 9, n_src * 9))
    for prs, dofs, corr, unc in blocks.values():
        B = corr if corrected else unc
        for k in range(dofs.shape[0]):
            i, j = dofs[k]
            M[9 * i:9 * i + 9, 9 * j:9 * j + 9] += B[k]
    return M


def interior_pairs(K, qx, qw, obs_pts, obs_ns, pts, tris, pairs, check0, finite_part, params):
    """Point-observer pair integrals (interior_pairs, tectosaur/assemble.cl
    :140-186): [n, 3, 3, 3] = [d_obs, b_src, d_src]; pairs [n, 3] =
    (obs point, src triangle, src clicks)."""
    qx, qw, obs_pts, obs_ns, pts, params = (_f64(a) for a in (qx, qw, obs_pts, obs_ns, pts,
                                                               params))
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    pairs = np.ascontiguousarray(np.asarray(pairs).reshape((-1, 3)), dtype=np.int64)
    n = pairs.shape[0]
    out = np.zeros(n * 27)
    if n > 0:
        lib().or_interior_pairs(KERNEL_IDS[K], qw.shape[0], _p(qx), _p(qw), _p(obs_pts),
                                _p(obs_ns), _p(pts), _p(tris, _i32), n, _p(pairs, _i64),
                                int(bool(check0)), int(bool(finite_part)), _p(params), _p(out))
    return out.reshape((n, 3, 3, 3))


def interior_pairs_sinh(K, t_rule, r_rule, obs_pts, obs_ns, pts, tris, pairs, params):
    """Point-triangle integrals by the polar sinh rule (t_rule, r_rule: Gauss
    rules on [0, 1] for the angle and the sinh variable): [n, 3, 3, 3],
    physical weights."""
    tq, tw = (_f64(a) for a in t_rule)
    rq, rw = (_f64(a) for a in r_rule)
    obs_pts, obs_ns, pts, params = (_f64(a) for a in (obs_pts, obs_ns, pts, params))
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    pairs = np.ascontiguousarray(np.asarray(pairs).reshape((-1, 3)), dtype=np.int64)
    n = pairs.shape[0]
    out = np.zeros(n * 27)
    if n > 0:
        lib().or_interior_pairs_sinh(KERNEL_IDS[K], tq.shape[0], _p(tq), _p(tw), rq.shape[0],
                                     _p(rq), _p(rw), _p(obs_pts), _p(obs_ns), _p(pts),
                                     _p(tris, _i32), n, _p(pairs, _i64), _p(params), _p(out))
    return out.reshape((n, 3, 3, 3))


def corner_radial(a, rho_max, n):
    """The radial integrals of the interior-corner rule at one angle
    (tectosaur/interior_corners.cl:79-83): (I1, I2, I3[3, 3], I4[3, 3]) with
    I4 the reference's value (zero: C integer division)."""
    a, n = _f64(a), _f64(n)
    out = np.zeros(20)
    lib().or_corner_radial(_p(a), float(rho_max), _p(n), _p(out))
    return out[0], out[1], out[2:11].reshape(3, 3), out[11:20].reshape(3, 3)


def interior_corners(qt, qw, obs_ns, pts, tris, pairs, params):
    """Observation point at a vertex of the source triangle
    (tectosaur/interior_corners.cl:23-175): [n, 3, 3, 3]."""
    qt, qw, obs_ns, pts, params = (_f64(a) for a in (qt, qw, obs_ns, pts, params))
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    pairs = np.ascontiguousarray(np.asarray(pairs).reshape((-1, 3)), dtype=np.int64)
    n = pairs.shape[0]
    out = np.zeros(n * 27)
    if n > 0:
        lib().or_interior_corners(qw.shape[0], _p(qt), _p(qw), _p(obs_ns), _p(pts),
                                  _p(tris, _i32), n, _p(pairs, _i64), _p(params), _p(out))
    return out.reshape((n, 3, 3, 3))


def near_pt_pairs(obs_pts, pts, tris, threshold):
    """(obs point, triangle) pairs with |x - c|^2 <= (r thr)^2, brute force
    (get_nearfield with zero observation radius, tectosaur/interior.py:33-37)."""
    cb, rb = tri_centroids_rs(pts, tris)
    ca = _f64(obs_pts)
    ra = np.zeros(ca.shape[0])
    cb, rb = _f64(cb), _f64(rb)
    cap = 1 << 16
    while True:
        out = np.zeros(2 * cap, dtype=np.int64)
        k = lib().or_near_pairs_bruteforce(ca.shape[0], _p(ca), _p(ra), cb.shape[0], _p(cb),
                                           _p(rb), float(threshold), _p(out, _i64), cap)
        if k <= cap:
            return out[:2 * k].reshape((k, 2))
        cap = k


def split_vertex_nearfield(pairs, obs_pts, pts, tris):
    """tectosaur/mesh/fast_find_nearfield.cpp:184-215: (vertex pairs [n,3] =
    (obs, tri, last vertex d within 1e-9 of the point), other pairs [n,3] =
    (obs, tri, 0)), each in input order."""
    vert, near = [], []
    for i, j in pairs:
        isv = -1
        for d in range(3):
            if np.linalg.norm(obs_pts[i] - pts[tris[j, d]]) < 1e-9:
                isv = d
        (vert if isv >= 0 else near).append((i, j, max(isv, 0)))
    return (np.array(vert, dtype=np.int64).reshape((-1, 3)),
            np.array(near, dtype=np.int64).reshape((-1, 3)))


def pairs_matvec(pairs, entries, v, n_obs):
    """y[3 obs + d] += sum entries[p, d, k] v[9 tri + k] (make_pairs_mat of
    tectosaur/interior.py:14-26 applied to v)."""
    y = np.zeros(n_obs * 3)
    e = np.asarray(entries).reshape((-1, 3, 9))
    for p, (i, j) in enumerate(np.asarray(pairs)[:, :2]):
        y[3 * i:3 * i + 3] += e[p] @ v[9 * j:9 * j + 9]
    return y


def interior_op_dot(obs_pts, obs_ns, pts, tris, K, v, params, far_rule, near_rule,
                    vertex_rule, threshold=4.0, sinh_rules=None):
    """InteriorOp.dot of tectosaur/interior.py:29-119, dense on the CPU:
    far (every pair, far rule, r2 == 0 skipped) + near pairs (near rule)
    - every close pair (far rule) + vertex pairs (corner rule).  The rules
    are (points, weights) so this module does not import the product;
    sinh_rules = (angle rule, sinh rule) on [0, 1] integrates the near pairs
    with the polar sinh rule instead."""
    obs_pts, obs_ns, pts, v = (_f64(a) for a in (obs_pts, obs_ns, pts, v))
    n_obs, n_tris = obs_pts.shape[0], tris.shape[0]
    allp = np.stack(np.meshgrid(np.arange(n_obs), np.arange(n_tris), indexing='ij'),
                    axis=2).reshape((-1, 2))
    allp = np.concatenate([allp, np.zeros((allp.shape[0], 1), np.int64)], axis=1)
    far = interior_pairs(K, far_rule[0], far_rule[1], obs_pts, obs_ns, pts, tris, allp, True,
                         False, params)
    y = pairs_matvec(allp, far, v, n_obs)
    close = near_pt_pairs(obs_pts, pts, tris, threshold)
    vert, near = split_vertex_nearfield(close, obs_pts, pts, tris)
    every = np.concatenate([near, vert], axis=0)
    every[:, 2] = 0
    if sinh_rules is None:
        near_e = interior_pairs(K, near_rule[0], near_rule[1], obs_pts, obs_ns, pts, tris, near,
                                False, False, params)
    else:
        near_e = interior_pairs_sinh(K, sinh_rules[0], sinh_rules[1], obs_pts, obs_ns, pts, tris,
                                     near, params)
    y += pairs_matvec(near, near_e, v, n_obs)
    y -= pairs_matvec(every, interior_pairs(K, far_rule[0], far_rule[1], obs_pts, obs_ns, pts,
                                            tris, every, False, False, params), v, n_obs)
    y += pairs_matvec(vert, interior_corners(vertex_rule[0], vertex_rule[1], obs_ns, pts, tris,
                                             vert, params), v, n_obs)
    return y
