"""Operator combinators (tectosaur/ops/sum_op.py, neg_op.py, composite_op.py,
dense_op.py): the reference's own tests (tests/test_op_combinators.py:6-27)
plus the device product ``dot_dev`` on torch tensors (CPU tensors here)."""
import numpy as np
import pytest
import torch

from tectosaur_amd.combinators import CompositeOp, DenseOp, MultOp, NegOp, SumOp


def test_composite_op():
    n, split_n = 20, 10
    rng = np.random.default_rng(0)
    mat = rng.random((n, n))
    v = rng.random(n)
    comp_op = CompositeOp(
        (DenseOp(mat[:split_n, :split_n]), 0, 0),
        (DenseOp(mat[split_n:, :split_n]), split_n, 0),
        (DenseOp(mat[:split_n, split_n:]), 0, split_n),
        (DenseOp(mat[split_n:, split_n:]), split_n, split_n))
    np.testing.assert_allclose(comp_op.dot(v), mat.dot(v), rtol=1e-13)
    for name in ('nearfield_dot', 'nearfield_no_correction_dot'):
        np.testing.assert_allclose(getattr(comp_op, name)(v), mat.dot(v), rtol=1e-13)
    assert np.all(comp_op.farfield_dot(v) == 0)
    out = torch.full((n,), np.nan, dtype=torch.float64)
    comp_op.dot_dev(torch.as_tensor(v), out)
    np.testing.assert_allclose(out.numpy(), mat.dot(v), rtol=1e-13)


def test_composite_op_overlap_and_shape():
    rng = np.random.default_rng(1)
    a, b = rng.random((4, 6)), rng.random((4, 6))
    v = rng.random(8)
    op = CompositeOp((DenseOp(a), 0, 0), (DenseOp(b), 2, 2), shape=(7, 8))
    want = np.zeros(7)
    want[0:4] += a @ v[0:6]
    want[2:6] += b @ v[2:8]
    assert op.shape == (7, 8)
    np.testing.assert_allclose(op.dot(v), want, rtol=1e-13)
    np.testing.assert_allclose(op.dot_dev(torch.as_tensor(v)).numpy(), want, rtol=1e-13)
    with pytest.raises(ValueError):
        CompositeOp((DenseOp(a), 0, 0), shape=(3, 6))


def test_sum_op():
    rng = np.random.default_rng(2)
    mat = rng.random((2, 2))
    v = rng.random(2)
    sum_op = SumOp([DenseOp(mat), DenseOp(mat)])
    np.testing.assert_allclose(sum_op.dot(v), 2 * mat.dot(v), rtol=1e-14)
    np.testing.assert_allclose(sum_op.dot_dev(torch.as_tensor(v)).numpy(), 2 * mat.dot(v),
                               rtol=1e-14)
    with pytest.raises(ValueError):
        SumOp([DenseOp(mat), DenseOp(np.ones((3, 3)))])


def test_mult_and_neg_op():
    rng = np.random.default_rng(3)
    mat = rng.random((5, 5))
    v = rng.random(5)
    for op, f in ((MultOp(DenseOp(mat), 2.5), 2.5), (NegOp(DenseOp(mat)), -1.0)):
        np.testing.assert_allclose(op.dot(v), f * mat.dot(v), rtol=1e-14)
        np.testing.assert_allclose(op.nearfield_dot(v), f * mat.dot(v), rtol=1e-14)
        np.testing.assert_allclose(op.dot_dev(torch.as_tensor(v)).numpy(), f * mat.dot(v),
                                   rtol=1e-14)
    # nesting: -(A + A) as a block of a composite
    nest = CompositeOp((NegOp(SumOp([DenseOp(mat), DenseOp(mat)])), 1, 0))
    np.testing.assert_allclose(nest.dot(v)[1:], -2 * mat.dot(v), rtol=1e-14)
    np.testing.assert_allclose(nest.dot_dev(torch.as_tensor(v)).numpy()[1:], -2 * mat.dot(v),
                               rtol=1e-14)


def test_package_exports():
    import tectosaur_amd as T
    for name in ('SumOp', 'MultOp', 'NegOp', 'CompositeOp', 'RegularizedSparseIntegralOp',
                 'RegularizedDenseIntegralOp', 'TriToTriDirectFarfieldOp', 'FMMFarfieldOp',
                 'MassOp', 'InteriorOp', 'make_rect', 'build_constraint_matrix'):
        assert callable(getattr(T, name))
    with pytest.raises(AttributeError):
        T.no_such_op
