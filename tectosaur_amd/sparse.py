"""Block-sparse matrices whose matvec runs on the device.

Mirrors tectosaur/util/sparse.py (``BCOOMatrix`` :14-42, ``BSRMatrix``
:44-70).  Data live in device memory; ``dot`` accepts host arrays (the
reference API) and ``dot_dev`` device tensors.  The product is the FP64
9x9 BSR kernel of sparse.hip (the reference's bcoomv9 / bsrmv9 run on the
host with OpenMP atomics); BCOO input is sorted by block row once on the
device so the matvec needs no atomics.
"""
import numpy as np
import torch

from . import _native as N


class BSRMatrix:
    """9x9-block CSR on the device: indptr [n_block_rows+1], indices [nnzb]
    (int32), data [nnzb, 81] float64."""

    def __init__(self, indptr, indices, data, shape):
        dev = N.require_device()
        self.device = dev
        self.indptr = N.upload(indptr, torch.int32, dev).contiguous()
        self.indices = N.upload(indices, torch.int32, dev).contiguous()
        d = N.upload(data, torch.float64, dev)
        self.data = d.reshape((-1, 81)).contiguous()
        self.data32 = None
        self.shape = tuple(shape)
        if self.shape[0] % 9 or self.shape[1] % 9:
            raise ValueError('shape must be a multiple of the 9x9 block size')
        if self.indptr.shape[0] != self.shape[0] // 9 + 1:
            raise ValueError('indptr length does not match the block rows')

    @property
    def blocksize(self):
        return 9

    @property
    def nnzb(self):
        return self.indices.shape[0]

    def make_f32(self):
        """Keep an FP32 copy of the block data for inexact (inner) products."""
        if self.data32 is None:
            self.data32 = self.data.to(torch.float32)
        return self

    def dot_dev(self, x, out=None, accumulate=False, prec='f64'):
        """y (+)= A x.  prec 'f64' uses the FP64 blocks; any other precision
        uses the FP32 copy when one was made (FP64 accumulation either way)."""
        if out is None:
            out = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
            accumulate = False
        if prec != 'f64' and self.data32 is not None:
            N.call('tct_bsr9_spmv_f32', self.shape[0] // 9, N.ptr(self.indptr),
                   N.ptr(self.indices), N.ptr(self.data32), N.ptr(x), N.ptr(out),
                   int(bool(accumulate)), N.stream_ptr())
        else:
            N.call('tct_bsr9_spmv', self.shape[0] // 9, N.ptr(self.indptr), N.ptr(self.indices),
                   N.ptr(self.data), N.ptr(x), N.ptr(out), int(bool(accumulate)),
                   N.stream_ptr())
        return out

    def dot(self, v):
        x = torch.as_tensor(np.ascontiguousarray(v, dtype=np.float64), device=self.device)
        return self.dot_dev(x).cpu().numpy()

    def to_scipy(self):
        import scipy.sparse
        data = self.data.cpu().numpy().reshape((-1, 9, 9))
        return scipy.sparse.bsr_matrix((data, self.indices.cpu().numpy(),
                                        self.indptr.cpu().numpy()), self.shape)

    def to_dense(self):
        return self.to_scipy().toarray()

    def to_dense_dev(self):
        """Dense FP64 device tensor of the matrix (blocks summed per slot)."""
        nbr, nbc = self.shape[0] // 9, self.shape[1] // 9
        counts = (self.indptr[1:] - self.indptr[:-1]).to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(nbr, device=self.device), counts)
        out = torch.zeros((nbr, nbc, 9, 9), dtype=torch.float64, device=self.device)
        out.index_put_((rows, self.indices.to(torch.int64)), self.data.reshape((-1, 9, 9)),
                       accumulate=True)
        return out.permute(0, 2, 1, 3).reshape(self.shape)

    @property
    def algorithmic_bytes(self):
        """HBM bytes one matvec must move: block data and column index per
        block, the row pointers, x read once, y written once."""
        return self.nnzb * (648 + 4) + self.indptr.shape[0] * 4 + 2 * 8 * self.shape[0]


def bcoo_to_bsr_dev(rows, cols, data, shape):
    """Sort device BCOO blocks by (row, col) and build the BSR matrix."""
    dev = data.device
    rows = rows.to(torch.int64)
    cols = cols.to(torch.int64)
    nbr = shape[0] // 9
    key = rows * (shape[1] // 9 + 1) + cols
    order = torch.argsort(key, stable=True)
    counts = torch.bincount(rows, minlength=nbr)
    indptr = torch.zeros(nbr + 1, dtype=torch.int64, device=dev)
    indptr[1:] = torch.cumsum(counts, 0)
    return BSRMatrix(indptr, cols[order], data.reshape((-1, 81))[order], shape)


class BCOOMatrix:
    """Block COO: rows, cols [n] and data [n, 9, 9]."""

    def __init__(self, rows, cols, data, shape):
        dev = N.require_device()
        self.rows = N.upload(rows, torch.int64, dev)
        self.cols = N.upload(cols, torch.int64, dev)
        self.data = N.upload(data, torch.float64, dev).reshape((-1, 81))
        self.shape = tuple(shape)
        self._bsr = None

    @property
    def blocksize(self):
        return 9

    def to_bsr(self):
        if self._bsr is None:
            self._bsr = bcoo_to_bsr_dev(self.rows, self.cols, self.data, self.shape)
        return self._bsr

    def dot(self, v):
        return self.to_bsr().dot(v)

    def dot_dev(self, x, out=None, accumulate=False):
        return self.to_bsr().dot_dev(x, out, accumulate)

    def to_dense(self):
        return self.to_bsr().to_dense()
