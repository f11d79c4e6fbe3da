"""Constraint matrix builder (host C++ through the C ABI), MassOp and the
continuity / free-edge builders: the exact assertions of the reference's
tests/test_constraints.py, plus continuity of cm-mapped fields."""
import numpy as np

from tectosaur_amd import mesh
from tectosaur_amd.constraints import (ConstraintEQ, Term, build_constraint_matrix,
                                       continuity_constraints, find_free_edges,
                                       free_edge_constraints)


def test_constraint_matrix():
    cm, rhs, _ = build_constraint_matrix([ConstraintEQ([Term(1, 0), Term(-1, 1)], 0.0)], 3)
    assert cm.shape == (3, 2)
    np.testing.assert_almost_equal(cm.todense(), [[1, 0], [1, 0], [0, 1]])
    np.testing.assert_almost_equal(rhs, 0)


def test_constraint_matrix_harder():
    cs = [ConstraintEQ([Term(1, 5), Term(-1, 1)], 0.0),
          ConstraintEQ([Term(1, 3), Term(0.25, 0)], 0.0),
          ConstraintEQ([Term(1, 2), Term(0.5, 3), Term(0.5, 4)], 0.0)]
    cm, rhs, _ = build_constraint_matrix(cs, 7)
    assert cm.shape == (7, 4)
    correct = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [-0.25, 0, 0, 0], [0.25, 0, -2, 0],
               [0, 1, 0, 0], [0, 0, 0, 1]]
    np.testing.assert_almost_equal(cm.todense(), correct)
    np.testing.assert_almost_equal(rhs, 0)


def test_constraint_matrix_rhs():
    _, rhs, _ = build_constraint_matrix([ConstraintEQ([Term(1, 0)], 2.0)], 1)
    np.testing.assert_almost_equal(rhs, [2.0])
    cs = [ConstraintEQ([Term(1, 0), Term(2, 2)], 2.0), ConstraintEQ([Term(1, 1), Term(1, 0)], 3.0)]
    _, rhs, _ = build_constraint_matrix(cs, 3)
    np.testing.assert_almost_equal(rhs, [0, 3.0, 1.0])
    cs = [ConstraintEQ([Term(1, 5), Term(-1, 1)], 0.0),
          ConstraintEQ([Term(1, 3), Term(0.25, 0)], 1.0),
          ConstraintEQ([Term(1, 2), Term(0.5, 3), Term(0.5, 4)], 2.0)]
    _, rhs, _ = build_constraint_matrix(cs, 7)
    np.testing.assert_almost_equal(rhs, [0, 0, 0, 1.0, 3.0, 0, 0])


def test_constraint_double_and_triple():
    cs = [ConstraintEQ([Term(1, 0), Term(1, 1), Term(1, 2)], 0.0),
          ConstraintEQ([Term(1, 0), Term(-1, 1), Term(1, 2)], 0.0)]
    cm, _, _ = build_constraint_matrix(cs, 3)
    np.testing.assert_almost_equal(cm.todense(), np.array([[1, 0, -1]]).T)
    cs = [ConstraintEQ([Term(1, 0), Term(1, 1), Term(1, 2), Term(1, 3)], 0.0),
          ConstraintEQ([Term(1, 0), Term(-1, 1), Term(1, 2), Term(1, 3)], 0.0),
          ConstraintEQ([Term(1, 0), Term(1, 1), Term(-1, 2), Term(1, 3)], 0.0)]
    cm, _, _ = build_constraint_matrix(cs, 4)
    np.testing.assert_almost_equal(cm.todense(), np.array([[1, 0, 0, -1]]).T)


def test_find_free_edges():
    free_es = find_free_edges(np.array([[0, 1, 2], [2, 1, 3]]))
    assert len(free_es) == 4
    for e in [(0, 0), (0, 2), (1, 1), (1, 2)]:
        assert e in free_es


def rect(n):
    return mesh.make_rect(n, n, [[-1.0, -1.0, 0], [-1.0, 1.0, 0], [1.0, 1.0, 0],
                                 [1.0, -1.0, 0]])


def test_free_edge_constraints():
    m = rect(3)
    dofs = [c.terms[0].dof for c in free_edge_constraints(m[1])]
    tri_pts = m[0][m[1]].reshape((-1, 3))
    near_origin = np.all(np.abs(tri_pts[:, :2]) < 0.1, axis=1)
    idx = np.where(~near_origin)[0]
    correct = set((np.tile(idx * 3, (3, 1)) + np.arange(3)[:, None]).reshape(-1).tolist())
    assert correct == set(dofs)
    assert len(dofs) == 18 * 3


def test_redundant_continuity():
    n = 13
    m = rect(n)
    cs = continuity_constraints(m[0], m[1], m[1].shape[0])
    cm, _, _ = build_constraint_matrix(cs, m[1].size * 3)
    assert cm.shape[1] == 3 * n ** 2


def test_cm_fields_are_continuous():
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, 2)
    cs = continuity_constraints(pts, tris, tris.shape[0])
    cm, _, _ = build_constraint_matrix(cs, tris.shape[0] * 9)
    assert cm.shape[1] == 3 * pts.shape[0]
    f = cm @ np.random.default_rng(1).normal(size=cm.shape[1])
    vals = f.reshape((-1, 3, 3))
    for v in range(pts.shape[0]):
        ti, ci = np.where(tris == v)
        assert np.allclose(vals[ti, ci], vals[ti[0], ci[0]])


def test_mass_op_integrates_constants():
    # sum of all entries of the scalar mass matrix = total area
    from tectosaur_amd.mass_op import MassOp
    pts, tris = rect(5)
    M = MassOp(3, pts, tris, tensor_dim=1)
    assert M.shape == (tris.shape[0] * 3,) * 2
    np.testing.assert_allclose(M.mat.sum(), 4.0, rtol=1e-14)
    M3 = MassOp(3, pts, tris)
    ones = np.ones(M3.shape[1])
    np.testing.assert_allclose(M3.dot(ones).sum(), 3 * 4.0, rtol=1e-14)


def test_continuity_cm_equals_general_builder():
    # the direct vertex-grouping cm equals build_constraint_matrix on
    # continuity + zero free-edge constraints
    from tectosaur_amd.traction_slip import continuity_cm, fault_mesh, FaultRequest
    meshes = [rect(6), mesh.make_icosphere([0, 0, 0], 1.0, 1),
              fault_mesh(FaultRequest(0, 30.0, 60.0, 1e4, 2.0, 300, np.zeros(3), 0))]
    for pts, tris in meshes:
        cs = continuity_constraints(pts, tris, tris.shape[0]) + free_edge_constraints(tris)
        cm, rhs, _ = build_constraint_matrix(cs, tris.shape[0] * 9)
        fast = continuity_cm(tris)
        assert fast.shape == cm.shape
        assert abs(fast - cm).max() == 0
        np.testing.assert_array_equal(rhs, 0)
