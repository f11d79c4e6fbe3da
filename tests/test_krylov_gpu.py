"""Krylov-basis kernels of the device GMRES (krylov.hip) against torch
FP64 on the same inputs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('m,n', [(1, 1000), (7, 5000), (40, 123457), (300, 2049), (60, 400000),
                                 (5, 1500001)])
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


def test_tick_gather_joins_streams_and_concatenates_segments():
    # segments written on two side streams are gathered on the current
    # stream after it joins both, copied to pinned host memory, and the done
    # event marks the copy; more segments than one launch carries (32)
    from tectosaur_amd import _native as N
    rng = np.random.default_rng(3)
    sides = [torch.cuda.Stream(), torch.cuda.Stream()]
    joins = [torch.cuda.Event() for _ in sides]
    for st, ev in zip(sides, joins):
        ev.record(st)   # creates the event
    lens = [int(x) for x in rng.integers(0, 3000, 45)] + [0, 70001]
    host_src = [rng.standard_normal(n) for n in lens]
    dev = []
    for i, a in enumerate(host_src):
        with torch.cuda.stream(sides[i % 2]):
            d = torch.empty(len(a), dtype=torch.float64, device='cuda')
            d.copy_(torch.as_tensor(a))
            d.mul_(2.0)   # work queued on the side stream, not yet joined
            dev.append(d)
    total = sum(lens)
    staging = torch.empty(total + 5, dtype=torch.float64, device='cuda')
    host = torch.empty(total, dtype=torch.float64, pin_memory=True)
    src = np.array([t.data_ptr() for t in dev], dtype=np.uint64)
    ln = np.array(lens, dtype=np.int32)
    st_arr = np.array([s.cuda_stream for s in sides], dtype=np.uint64)
    ev_arr = np.array([e.cuda_event for e in joins], dtype=np.uint64)
    done = torch.cuda.Event()
    done.record()
    N.call('tct_tick_gather', len(sides), st_arr.ctypes.data, ev_arr.ctypes.data, len(lens),
           src.ctypes.data, ln.ctypes.data, N.ptr(staging), staging.shape[0], N.ptr(host),
           N.stream_ptr(), done.cuda_event)
    done.synchronize()
    ref = 2.0 * np.concatenate(host_src)
    assert np.array_equal(host.numpy(), ref)
    # a staging buffer one value short is refused before anything runs
    with pytest.raises(N.TctError, match='staging'):
        N.call('tct_tick_gather', len(sides), st_arr.ctypes.data, ev_arr.ctypes.data, len(lens),
               src.ctypes.data, ln.ctypes.data, N.ptr(staging), total - 1, N.ptr(host),
               N.stream_ptr(), None)
    torch.cuda.synchronize()


def test_arnoldi_step_is_graph_replay_then_cgs2():
    # the fused step equals replaying the captured matvec and calling cgs2
    from tectosaur_amd import _native as N
    m, n = 5, 4099
    rng = np.random.default_rng(4)
    q, _ = np.linalg.qr(rng.standard_normal((n, m)))
    V = torch.as_tensor(np.ascontiguousarray(q.T), device='cuda')
    A = torch.as_tensor(rng.standard_normal(n), device='cuda')
    x = torch.as_tensor(rng.standard_normal(n), device='cuda')
    w = torch.empty(n, dtype=torch.float64, device='cuda')
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        torch.mul(A, x, out=w)   # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        torch.mul(A, x, out=w)
    nd = ctypes.c_int64(0)
    N.call('tct_krylov_scratch', m, n, ctypes.byref(nd))
    scratch = torch.empty(nd.value, dtype=torch.float64, device='cuda')
    h = torch.empty(m + 1, dtype=torch.float64, device='cuda')
    vn = torch.empty(n, dtype=torch.float64, device='cuda')
    w.zero_()
    torch.cuda.synchronize()
    N.call('tct_arnoldi_step', g.raw_cuda_graph_exec(), m, n, n, N.ptr(V), N.ptr(w), N.ptr(h),
           N.ptr(scratch), nd.value, N.ptr(vn), None, N.stream_ptr())
    torch.cuda.synchronize()
    w0 = A * x
    h1 = V @ w0
    wr = w0 - V.t() @ h1
    h2 = V @ wr
    wr = wr - V.t() @ h2
    ref_h = torch.cat([h1 + h2, torch.linalg.vector_norm(wr).reshape(1)])
    assert torch.allclose(h, ref_h, rtol=1e-12, atol=1e-12 * float(ref_h.abs().max()))
    assert torch.allclose(w, wr, rtol=0, atol=1e-12 * float(w0.abs().max()))
    assert torch.equal(vn, w / h[m])
    with pytest.raises(N.TctError, match='null graph'):
        N.call('tct_arnoldi_step', None, m, n, n, N.ptr(V), N.ptr(w), N.ptr(h), N.ptr(scratch),
               nd.value, None, None, N.stream_ptr())
