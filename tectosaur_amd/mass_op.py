"""``MassOp``: the P1 Galerkin mass matrix, drop-in for
tectosaur/ops/mass_op.py:9-42 (with the entries of
tectosaur/ops/_mass_op.cpp:11-40).  Built once on the host (block diagonal,
27 entries per triangle for tensor_dim 3), applied on the device by the CSR
kernel of sparse.hip."""
import numpy as np
import scipy.sparse
import torch

from . import _native as N
from .quadrature import gauss2d_tri


class DeviceCSR:
    """FP64 CSR matrix on the device (rows < 2^31)."""

    def __init__(self, A):
        A = scipy.sparse.csr_matrix(A)
        A.sort_indices()
        dev = N.require_device()
        self.device = dev
        self.shape = A.shape
        self.indptr = N.upload(A.indptr.astype(np.int32), None, dev)
        self.indices = N.upload(A.indices.astype(np.int32), None, dev)
        self.data = N.upload(A.data.astype(np.float64), None, dev)

    def dot_dev(self, x, out=None, accumulate=False):
        if out is None:
            out = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
            accumulate = False
        N.call('tct_csr_spmv', self.shape[0], N.ptr(self.indptr), N.ptr(self.indices),
               N.ptr(self.data), N.ptr(x), N.ptr(out), int(bool(accumulate)), N.stream_ptr())
        return out


class MassOp:
    def __init__(self, nq, pts, tris, tensor_dim=3):
        qx, qw = gauss2d_tri(nq)
        pts = np.asarray(pts, dtype=np.float64)
        tris = np.asarray(tris)
        tri_pts = pts[tris]
        basis = np.stack([1 - qx[:, 0] - qx[:, 1], qx[:, 0], qx[:, 1]], axis=1)
        un = np.cross(tri_pts[:, 2] - tri_pts[:, 0], tri_pts[:, 2] - tri_pts[:, 1])
        jac = np.linalg.norm(un, axis=1)
        bf = np.einsum('q,qa,qb->ab', qw, basis, basis).reshape(-1)
        n = tris.shape[0]
        td = tensor_dim
        n_rows = n * 3 * td
        self.shape = (n_rows, n_rows)
        # row (tri i, basis b1, component d) holds the three entries
        # jac_i bf[b1][b2] at columns (i, b2, d), b2 = 0, 1, 2 (already in
        # column order): the CSR arrays are formed directly, no sort
        r = np.arange(n_rows)
        i = r // (3 * td)
        b1 = (r // td) % 3
        d = r % td
        indices = (3 * i[:, None] + np.arange(3)[None, :]) * td + d[:, None]
        data = jac[i][:, None] * bf.reshape(3, 3)[b1]
        indptr = np.arange(0, 3 * n_rows + 1, 3)
        self.mat = scipy.sparse.csr_matrix((data.reshape(-1), indices.reshape(-1), indptr),
                                           shape=self.shape)
        self._dev = None

    @property
    def dev(self):
        if self._dev is None:
            self._dev = DeviceCSR(self.mat)
        return self._dev

    def dot(self, v):
        return self.mat.dot(v)

    def dot_dev(self, x, out=None, accumulate=False):
        return self.dev.dot_dev(x, out, accumulate)

    def nearfield_dot(self, v):
        return self.dot(v)

    def nearfield_no_correction_dot(self, v):
        return self.dot(v)

    def farfield_dot(self, v):
        shape = [self.shape[0]]
        shape.extend(v.shape[1:])
        return np.zeros(shape)
