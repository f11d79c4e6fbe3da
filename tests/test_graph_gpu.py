"""HIP-graph capture of the constrained operator: every replay, with eager
device work and synchronisations between replays, must equal the eager
apply bit for bit (same kernels on the same buffers)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ORD = dict(nq_coincident=4, nq_edge_adj=4, nq_vert_adjacent=3, nq_far=1, nq_near=3,
           near_threshold=2.5, quad_mass_order=3)


def test_captured_apply_equals_eager_on_every_replay():
    from tectosaur_amd.solver import ConstrainedOperator
    from tectosaur_amd.traction_slip import (FaultRequest, TractionSlipProblem, fault_mesh,
                                             prepare_problem, request_traction)
    req = FaultRequest(0, 35.0, 55.0, 8000.0, 1.5, 700, np.array([0.3, -0.7, 0.2]), 5)
    pts, tris = fault_mesh(req)
    prep = prepare_problem(pts, tris, request_traction(req, tris.shape[0]), orders=ORD,
                           tree=dict(order=3, mac=2.0, min_tris=0))
    prob = TractionSlipProblem(None, None, None, prepared=prep)
    op = ConstrainedOperator(prob.H, prob.cm)
    n = op.shape[0]
    rng = np.random.default_rng(0)
    xs = [torch.as_tensor(rng.normal(size=n), device='cuda') for _ in range(5)]
    precs = ('f64', 'f32', 'tree')
    eager = {p: [op.apply(x, None, p).clone() for x in xs] for p in precs}
    op.enable_graphs(precs)
    for rep in range(2):
        for p in precs:
            for i, x in enumerate(xs):
                y = op.apply(x, None, p)
                # eager work and a sync between replays
                z = (x * 2.0).sum()
                torch.cuda.synchronize()
                assert torch.equal(y, eager[p][i]), (rep, p, i)
                assert torch.isfinite(z)


def test_graphs_captured_in_separate_calls_stay_valid():
    # enable_graphs called once per precision: the graph input and output
    # buffers are shared, so the first graph still reads and writes them
    from tectosaur_amd.solver import ConstrainedOperator
    from tectosaur_amd.traction_slip import (FaultRequest, TractionSlipProblem, fault_mesh,
                                             prepare_problem, request_traction)
    req = FaultRequest(0, 120.0, 40.0, 6000.0, 2.0, 500, np.array([-0.4, 0.1, 0.6]), 2)
    pts, tris = fault_mesh(req)
    prep = prepare_problem(pts, tris, request_traction(req, tris.shape[0]), orders=ORD)
    prob = TractionSlipProblem(None, None, None, prepared=prep)
    op = ConstrainedOperator(prob.H, prob.cm)
    x = torch.as_tensor(np.random.default_rng(1).normal(size=op.shape[0]), device='cuda')
    eager = {p: op.apply(x, None, p).clone() for p in ('f32', 'f64')}
    op.enable_graphs(('f32',))
    gx = op.gx
    op.enable_graphs(('f64',))
    assert op.gx is gx
    for p in ('f32', 'f64', 'f32'):
        assert torch.equal(op.apply(x, None, p), eager[p]), p
        op.gx.copy_(x)
        op.replay(p)
        torch.cuda.synchronize()
        assert torch.equal(op.gy, eager[p]), p


def test_upload_from_builder_threads_keeps_every_copy():
    # builder threads upload host arrays through pinned memory on streams of
    # their own without waiting for the device, and drop each host array at
    # once; torch's pinned blocks are reused only after their copies finish,
    # so every device copy must equal its source
    import concurrent.futures
    from tectosaur_amd import _native as N

    def work(seed):
        st = torch.cuda.Stream()
        rng = np.random.default_rng(seed)
        outs, refs = [], []
        with torch.cuda.stream(st):
            for _ in range(40):
                a = rng.standard_normal(int(rng.integers(1, 150000)))
                refs.append(a.copy())
                outs.append(N.upload(a, torch.float64))
                del a
                outs[-1].mul_(1.0)   # queued device work behind the copy
        st.synchronize()
        return all(np.array_equal(o.cpu().numpy(), r) for o, r in zip(outs, refs))

    with concurrent.futures.ThreadPoolExecutor(4) as ex:
        assert all(ex.map(work, range(4)))
