"""tct_cgs2_noalloc (CGS2 with its partial sums in the library's static
device array) against tct_cgs2 with a caller workspace: bit-identical h, w
and next basis vector, on two calls in a row, the second smaller, and the
size limit refused before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _basis(m, n, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)))
    V = torch.as_tensor(np.ascontiguousarray(q.T), device='cuda')     # [m][n] orthonormal rows
    w0 = torch.as_tensor(rng.standard_normal(n), device='cuda')
    return V, w0


def _cgs2_with_workspace(N, V, w0):
    m, n = V.shape
    nd = ctypes.c_int64(0)
    N.call('tct_krylov_scratch', m, n, ctypes.byref(nd))
    scratch = torch.empty(nd.value, dtype=torch.float64, device='cuda')
    w = w0.clone()
    h = torch.empty(m + 1, dtype=torch.float64, device='cuda')
    vn = torch.empty(n, dtype=torch.float64, device='cuda')
    N.call('tct_cgs2', m, n, n, N.ptr(V), N.ptr(w), N.ptr(h), N.ptr(scratch), nd.value,
           N.ptr(vn), None, N.stream_ptr())
    torch.cuda.synchronize()
    return w, h, vn


def _cgs2_noalloc(N, V, w0):
    m, n = V.shape
    w = w0.clone()
    h = torch.empty(m + 1, dtype=torch.float64, device='cuda')
    vn = torch.empty(n, dtype=torch.float64, device='cuda')
    vn2 = torch.empty(n, dtype=torch.float64, device='cuda')
    N.call('tct_cgs2_noalloc', m, n, n, N.ptr(V), N.ptr(w), N.ptr(h), N.ptr(vn), N.ptr(vn2),
           N.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(vn2, vn)
    return w, h, vn


def test_cgs2_noalloc_equals_cgs2_on_two_calls():
    from tectosaur_amd import _native as N
    # the first call takes the many-block path (separate column sums), the
    # second is smaller and reuses the static array the first one filled
    for m, n, seed in ((64, 300000, 1), (7, 5000, 2)):
        V, w0 = _basis(m, n, seed)
        w_ref, h_ref, vn_ref = _cgs2_with_workspace(N, V, w0)
        w, h, vn = _cgs2_noalloc(N, V, w0)
        assert torch.equal(h, h_ref), (m, n)
        assert torch.equal(w, w_ref), (m, n)
        assert torch.equal(vn, vn_ref), (m, n)
        # and the projection itself against torch FP64
        h1 = V @ w0
        wr = w0 - V.t() @ h1
        h2 = V @ wr
        wr = wr - V.t() @ h2
        assert torch.allclose(h[:m], h1 + h2, rtol=1e-12, atol=1e-12 * float(w0.abs().max()))
        assert torch.allclose(w, wr, rtol=0, atol=1e-12 * float(w0.abs().max()))


def test_cgs2_noalloc_refuses_sizes_beyond_its_array():
    from tectosaur_amd import _native as N
    L = N.lib()
    limit = 1 << 22
    m, n = 3000, 2_000_000          # needs 2 * 1024 * 3000 + 6000 doubles
    nd = ctypes.c_int64(0)
    N.call('tct_krylov_scratch', m, n, ctypes.byref(nd))
    assert nd.value > limit
    buf = torch.zeros(8, dtype=torch.float64, device='cuda')
    p = N.ptr(buf)
    rc = L.tct_cgs2_noalloc(m, n, n, p, p, p, None, None, N.stream_ptr())
    assert rc < 0 and b'scratch' in L.tct_last_error()
