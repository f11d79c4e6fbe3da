"""Operator combinators: ``SumOp``, ``MultOp`` / ``NegOp``, ``CompositeOp``
and ``DenseOp``, drop-ins for tectosaur/ops/sum_op.py:1-16,
tectosaur/ops/neg_op.py:1-21, tectosaur/ops/composite_op.py:3-37 and
tectosaur/ops/dense_op.py:3-18.

They keep the reference's duck-typed protocol (``shape``, ``dot``,
``nearfield_dot``, ``nearfield_no_correction_dot``, ``farfield_dot`` on host
arrays) and add ``dot_dev(x, out=None, prec=None)`` on FP64 device tensors,
so a combination of device operators (for example the fault / surface
subset blocks a traction problem is built from) applies with no host round
trip.  ``prec`` is passed to the operators that take it.
"""
import inspect

import numpy as np
import torch


def _dot_dev(op, x, out, prec):
    """op.dot_dev(x, out) into ``out`` (overwritten), passing ``prec`` to
    operators whose dot_dev takes a precision (those without one, such as
    MassOp, are exact FP64 products).  A None ``prec`` keeps each operator's
    default."""
    if prec is not None and 'prec' in inspect.signature(op.dot_dev).parameters:
        y = op.dot_dev(x, out, prec=prec)
    else:
        y = op.dot_dev(x, out)
    if y is not out:
        out.copy_(y)
    return out


class SumOp:
    """sum_k ops[k] (tectosaur/ops/sum_op.py:1-16); every op has the shape
    of the first."""

    def __init__(self, ops):
        self.ops = list(ops)
        self.shape = self.ops[0].shape
        for op in self.ops[1:]:
            if tuple(op.shape) != tuple(self.shape):
                raise ValueError('SumOp operands differ in shape: %s vs %s' %
                                 (op.shape, self.shape))

    def nearfield_dot(self, v):
        return sum(op.nearfield_dot(v) for op in self.ops)

    def nearfield_no_correction_dot(self, v):
        return sum(op.nearfield_no_correction_dot(v) for op in self.ops)

    def dot(self, v):
        return sum(op.dot(v) for op in self.ops)

    def farfield_dot(self, v):
        return sum(op.farfield_dot(v) for op in self.ops)

    def dot_dev(self, x, out=None, prec=None):
        if out is None:
            out = torch.empty(self.shape[0], dtype=torch.float64, device=x.device)
        _dot_dev(self.ops[0], x, out, prec)
        if len(self.ops) > 1:
            tmp = torch.empty_like(out)
            for op in self.ops[1:]:
                out += _dot_dev(op, x, tmp, prec)
        return out


class MultOp:
    """factor * op (tectosaur/ops/neg_op.py:1-17)."""

    def __init__(self, op, factor):
        self.op = op
        self.shape = op.shape
        self.factor = factor

    def nearfield_dot(self, v):
        return self.factor * self.op.nearfield_dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.factor * self.op.nearfield_no_correction_dot(v)

    def dot(self, v):
        return self.factor * self.op.dot(v)

    def farfield_dot(self, v):
        return self.factor * self.op.farfield_dot(v)

    def dot_dev(self, x, out=None, prec=None):
        if out is None:
            out = torch.empty(self.shape[0], dtype=torch.float64, device=x.device)
        _dot_dev(self.op, x, out, prec)
        out.mul_(self.factor)
        return out


def NegOp(op):
    """-op (tectosaur/ops/neg_op.py:19-21)."""
    return MultOp(op, -1)


class CompositeOp:
    """Block operator: each (op, row_start, col_start) adds op applied to
    v[col_start : col_start + op.shape[1]] into rows [row_start, row_start +
    op.shape[0]) (tectosaur/ops/composite_op.py:3-37).  Blocks may overlap;
    their contributions add.  ``shape`` defaults to the bounding box."""

    def __init__(self, *ops_and_starts, shape=None):
        self.ops = [el[0] for el in ops_and_starts]
        self.row_start = [int(el[1]) for el in ops_and_starts]
        self.col_start = [int(el[2]) for el in ops_and_starts]
        n_rows = max(r + op.shape[0] for op, r in zip(self.ops, self.row_start))
        n_cols = max(c + op.shape[1] for op, c in zip(self.ops, self.col_start))
        self.shape = (n_rows, n_cols) if shape is None else tuple(shape)
        if self.shape[0] < n_rows or self.shape[1] < n_cols:
            raise ValueError('CompositeOp shape %s is smaller than its blocks (%d, %d)' %
                             (self.shape, n_rows, n_cols))

    def generic_dot(self, v, dot_name):
        v = np.asarray(v)
        out = np.zeros([self.shape[0]] + list(v.shape[1:]))
        for op, r0, c0 in zip(self.ops, self.row_start, self.col_start):
            out[r0:r0 + op.shape[0]] += getattr(op, dot_name)(v[c0:c0 + op.shape[1]])
        return out

    def nearfield_dot(self, v):
        return self.generic_dot(v, 'nearfield_dot')

    def nearfield_no_correction_dot(self, v):
        return self.generic_dot(v, 'nearfield_no_correction_dot')

    def dot(self, v):
        return self.generic_dot(v, 'dot')

    def farfield_dot(self, v):
        return self.generic_dot(v, 'farfield_dot')

    def dot_dev(self, x, out=None, prec=None):
        if out is None:
            out = torch.zeros(self.shape[0], dtype=torch.float64, device=x.device)
        else:
            out.zero_()
        tmp = {}
        for op, r0, c0 in zip(self.ops, self.row_start, self.col_start):
            n = op.shape[0]
            if n not in tmp:
                tmp[n] = torch.empty(n, dtype=torch.float64, device=x.device)
            xs = x[c0:c0 + op.shape[1]].contiguous()
            out[r0:r0 + n] += _dot_dev(op, xs, tmp[n], prec)
        return out


class DenseOp:
    """A dense matrix as an operator (tectosaur/ops/dense_op.py:3-18): its
    whole action counts as nearfield.  ``dot_dev`` multiplies a device copy
    made on first use."""

    def __init__(self, mat):
        self.mat = mat
        self.shape = mat.shape
        self._dev = None

    def dot(self, v):
        return self.mat.dot(v)

    def nearfield_dot(self, v):
        return self.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.dot(v)

    def farfield_dot(self, v):
        return np.zeros(self.shape[0])

    def dot_dev(self, x, out=None, prec=None):
        if self._dev is None or self._dev.device != x.device:
            self._dev = torch.as_tensor(np.ascontiguousarray(self.mat, dtype=np.float64),
                                        device=x.device)
        y = torch.mv(self._dev, x)
        if out is None:
            return y
        out.copy_(y)
        return out
