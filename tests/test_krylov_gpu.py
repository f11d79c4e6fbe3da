"""Krylov-basis kernels of the device GMRES (krylov.hip) against torch
FP64 on the same inputs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('m,n', [(1, 1000), (7, 5000), (40, 123457), (300, 2049)])
def test_cgs2_and_combine_match_torch(m, n):
    from tectosaur_amd import _native as N
    rng = np.random.default_rng(m)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)))
    V = torch.as_tensor(np.ascontiguousarray(q.T), device='cuda')     # [m][n] orthonormal rows
    w0 = torch.as_tensor(rng.standard_normal(n), device='cuda')
    nd = ctypes.c_int64(0)
    N.call('tct_krylov_scratch', m, n, ctypes.byref(nd))
    scratch = torch.empty(nd.value, dtype=torch.float64, device='cuda')
    h = torch.empty(m + 1, dtype=torch.float64, device='cuda')
    w = w0.clone()
    vn = torch.empty(n, dtype=torch.float64, device='cuda')
    vn2 = torch.empty(n, dtype=torch.float64, device='cuda')
    N.call('tct_cgs2', m, n, n, N.ptr(V), N.ptr(w), N.ptr(h), N.ptr(scratch), nd.value,
           N.ptr(vn), N.ptr(vn2), N.stream_ptr())
    h1 = V @ w0
    wr = w0 - V.t() @ h1
    h2 = V @ wr
    wr = wr - V.t() @ h2
    ref_h = torch.cat([h1 + h2, torch.linalg.vector_norm(wr).reshape(1)])
    assert torch.allclose(h, ref_h, rtol=1e-12, atol=1e-12 * float(ref_h.abs().max()))
    assert torch.allclose(w, wr, rtol=0, atol=1e-12 * float(w0.abs().max()))
    assert torch.equal(vn, w / h[m]) and torch.equal(vn2, vn)
    # repeated calls give bitwise identical results (fixed reduction order)
    w2 = w0.clone()
    h2b = torch.empty_like(h)
    N.call('tct_cgs2', m, n, n, N.ptr(V), N.ptr(w2), N.ptr(h2b), N.ptr(scratch), nd.value,
           None, None, N.stream_ptr())
    assert torch.equal(h2b, h) and torch.equal(w2, w)
    # combine: out = beta out + alpha V^T c
    c = torch.as_tensor(rng.standard_normal(m), device='cuda')
    out = torch.as_tensor(rng.standard_normal(n), device='cuda')
    ref = 0.5 * out + 2.0 * (V.t() @ c)
    N.call('tct_krylov_combine', m, n, n, N.ptr(V), N.ptr(c), 2.0, 0.5, N.ptr(out),
           N.stream_ptr())
    assert torch.allclose(out, ref, rtol=1e-12, atol=1e-12 * float(ref.abs().max()))
