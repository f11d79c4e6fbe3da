"""Benchmark: H-kernel farfield pair-evaluations per second on one MI355X.

Workload (the two-surface farfield configuration): a 40 km x 40 km free
surface of 256 x 256 squares (131,072 triangles) plus a 60-degree dipping
planar fault of 128 x 64 squares (16,384 triangles), concatenated with
duplicate points merged; 3-point triangle rule for observation and source
(442,368 quadrature points each); G = 3e10 Pa, nu = 0.25; slip N(0, 1) from
numpy default_rng(0).  One step = one farfield matvec of the summed
operator's farfield (gather, N-body, scatter) with the input resident in
HBM: 442,368^2 = 1.957e11 pair evaluations.

Printed JSON (one line): value = pair-evals/s of the FP32 H matvec; per
kernel FP32/FP64 rates and FP32-vs-FP64 errors; the roofline of the N-body
kernel from HIP events; the CPU oracle timed on a bounded row sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (datasheet)
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector peak (datasheet)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E bandwidth (datasheet)
PARAMS = [3e10, 0.25]
# Algorithmic FP64 flops of one RH3 quadrature-point evaluation in the pair
# integrator: two P1 maps (2 x 3 coords x 5) + D (3) + basis weights (4)
# + r^2 (5) + rsqrt (1) + w/r, w/r^3 (3) + 1/r moment (1) + six D_p D_t / r^3
# moments (6 x 3) = 65.
RH_NEAR_FLOPS_PER_QP = 65
# HBM bytes per launch from hardware counters (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE in separate runs, scaled by a 1 GiB copy measured in the same
# runs), read from the profile summary, which records the launch shape and
# a hash of the kernel sources it was measured on: a figure measured on
# other sources or another launch shape is not reported.
TRAFFIC_SOURCE = 'profiles/far_p2p_H_two_surface_442k_pmc_traffic.md'
TRAFFIC_KERNEL_SOURCES = {
    'far_p2p_f32_kernel<H>': ('tectosaur_amd/csrc/far_p2p.hip', 'tectosaur_amd/csrc/far_pair.h'),
    'bsr9_kernel': ('tectosaur_amd/csrc/sparse.hip',)}


def sources_hash(files):
    import hashlib
    h = hashlib.sha256()
    for f in files:
        with open(os.path.join(ROOT, f), 'rb') as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profiled_traffic(kernel, tile):
    """HBM bytes per launch of ``kernel`` at launch shape ``tile`` from the
    machine-readable line of the traffic profile, or (None, reason)."""
    import re
    try:
        with open(os.path.join(ROOT, TRAFFIC_SOURCE)) as f:
            txt = f.read()
    except OSError:
        return None, 'no profile'
    have = sources_hash(TRAFFIC_KERNEL_SOURCES[kernel])
    for m in re.finditer(r'<!-- traffic (\{.*?\}) -->', txt):
        e = json.loads(m.group(1))
        if e.get('kernel') != kernel or list(e.get('tile', [])) != list(tile):
            continue
        if e.get('sources') != have:
            return None, 'profiled on other kernel sources (%s, now %s)' % (e.get('sources'), have)
        return e['bytes'], TRAFFIC_SOURCE
    return None, 'no profile of this kernel at this launch shape'


def two_surface_mesh():
    from tectosaur_amd import mesh
    return mesh.two_surface_mesh()


def time_matvec(op, v_dev, prec, steps, warmup):
    """Seconds per matvec (wall clock over ``steps`` matvecs) and the mean
    duration of the N-body launch inside those same matvecs (HIP events on
    the stream it runs on)."""
    import torch
    out = torch.empty(op.shape[0], dtype=torch.float64, device=op.device)
    for _ in range(warmup):
        op.dot_dev(v_dev, out, prec)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in evs:
        op.dot_dev(v_dev, out, prec, nbody_events=e)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    kt = float(np.mean([a.elapsed_time(b) for a, b in evs])) * 1e-3
    return dt, out, kt


def stream_event_time(fn, reps):
    """Average wall time of ``fn()`` on torch's current stream, from HIP
    events recorded on that stream around each call."""
    import torch
    s = torch.cuda.current_stream()
    t = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        t.append(a.elapsed_time(b) * 1e-3)
    return float(np.mean(t))


def valu_calibration(blocks=8192, iters=4000, reps=2):
    """VALU issue rates of dependent-free FMA streams (tct_calib_valu):
    FP32 and FP64, every operand in a VGPR or one operand in an SGPR.  Each
    launch issues blocks * 4 waves * 32 * iters FMA wave-instructions.
    gfx950 issues two VALU instructions of one SIMD in a cycle only when
    neither reads an SGPR, so the FP32 VGPR rate is the chip's FP32 vector
    peak without packed instructions."""
    import torch
    from tectosaur_amd import _native as N
    out = torch.empty(blocks * 256, dtype=torch.float32, device='cuda')
    res = {'blocks': blocks, 'iters': iters}
    names = {0: 'f32_vgpr', 1: 'f32_sgpr', 2: 'f64_vgpr', 3: 'f64_sgpr'}
    for mode, name in names.items():
        N.call('tct_calib_valu', mode, blocks, iters, N.ptr(out), N.stream_ptr())
        dt = stream_event_time(
            lambda: N.call('tct_calib_valu', mode, blocks, iters, N.ptr(out), N.stream_ptr()),
            reps)
        wi = blocks * 4.0 * 32 * iters
        res[name] = {'ms': dt * 1e3, 'wave_instr_per_s': wi / dt,
                     'tflops': wi * 64 * 2 / dt / 1e12}
    print('# valu calibration: %s' % json.dumps(res), flush=True)
    return res


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size (reads nbytes, writes nbytes), run
    so that hardware byte counters (FETCH_SIZE, WRITE_SIZE) read in the same
    process can be scaled to bytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device='cuda')
    dst = torch.empty_like(src)
    dt = stream_event_time(lambda: dst.copy_(src), reps)
    del src, dst
    torch.cuda.empty_cache()
    return {'bytes_read': nbytes, 'bytes_written': nbytes, 'ms': dt * 1e3,
            'GB_per_s': 2 * nbytes / dt / 1e9}


def nearfield_bench(level=5, nq=8, thr=3.0, reps=20):
    """Nearfield on the icosphere workload: level-5 icosphere (20,480
    triangles, unit radius), every rule of order 8, near threshold 3, RH3
    near with the H3 farfield correction, FP64.  Reports the pair
    integrator's rate on every pair class (coincident, edge- and
    vertex-adjacent with their singular rules and the r = 0 test, and the
    plain-near class, the largest) and the BSR matvec (the per-iteration
    nearfield gather) against HBM bandwidth."""
    import torch
    from tectosaur_amd import mesh
    from tectosaur_amd.nearfield_op import RegularizedNearfieldIntegralOp
    from tectosaur_amd.pairs_integrator import PairsIntegrator
    pts, tris = mesh.make_icosphere([0, 0, 0], 1.0, level)
    n = tris.shape[0]
    idx = np.arange(n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op = RegularizedNearfieldIntegralOp(pts, tris, idx, idx, nq, nq, nq, 2, nq, thr, 'elasticRH3',
                                        'elasticH3', PARAMS, np.float32, keep_classes=True,
                                        keep_uncorrected=False)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    from tectosaur_amd import triangle_rules
    pi = PairsIntegrator('elasticRH3', PARAMS, np.float32, 2, nq, pts, tris)
    rules = {'co': (pi._rule(('co', nq), lambda: triangle_rules.coincident_quad(nq)), True),
             'ea': (pi._rule(('ea', nq), lambda: triangle_rules.edge_adj_quad(nq)), True),
             'va': (pi._rule(('va', nq, nq, nq),
                             lambda: triangle_rules.vertex_adj_quad(nq, nq, nq)), True),
             'near': (pi.near_q, False)}
    per_class = {}
    for name, (q, check0) in rules.items():
        prs = op.classes[name][0]
        p_dev = torch.as_tensor(np.ascontiguousarray(prs.astype(np.int32)), device='cuda')
        out = torch.empty((prs.shape[0], 81), dtype=torch.float64, device='cuda')
        nqp = q[0].shape[0]
        kt = stream_event_time(lambda: pi.pairs_quad_dev(q, p_dev, check0, out), 3)
        qps = prs.shape[0] * nqp / kt
        per_class[name] = {'pairs': int(prs.shape[0]), 'qp_per_pair': nqp, 'ms': kt * 1e3,
                           'qp_evals_per_s': qps,
                           'frac_fp64_peak': qps * RH_NEAR_FLOPS_PER_QP / 1e12 /
                           PEAK_FP64_TFLOPS}
        del out, p_dev
    op.classes = {}
    near = per_class['near']
    torch.cuda.empty_cache()
    mat = op.mat
    x = torch.as_tensor(np.random.default_rng(0).standard_normal(mat.shape[1]), device='cuda')
    y = torch.empty(mat.shape[0], dtype=torch.float64, device='cuda')
    mat.dot_dev(x, y)
    bt = stream_event_time(lambda: mat.dot_dev(x, y), reps)
    gbs = mat.algorithmic_bytes / bt / 1e9
    bsr_traffic, bsr_src = (profiled_traffic('bsr9_kernel', [int(mat.nnzb)])
                            if level == 5 and nq == 8 else (None, 'not profiled'))
    res = {'workload': 'icosphere level %d (%d tris), rules of order %d, near threshold %g, '
                       'RH3 near / H3 correction, FP64' % (level, n, nq, thr),
           'pair_counts': op.counts, 'build_s': t_build,
           'pairs_kernel': dict(near, **{'class': 'near',
                                         'flops_per_qp': RH_NEAR_FLOPS_PER_QP,
                                         'tflops': near['qp_evals_per_s'] *
                                         RH_NEAR_FLOPS_PER_QP / 1e12}),
           'pairs_kernel_by_class': per_class,
           'bsr_matvec': {'blocks': int(mat.nnzb), 'bytes': int(mat.algorithmic_bytes),
                          'ms': bt * 1e3, 'GB_per_s': gbs, 'frac_hbm_peak': gbs / PEAK_HBM_GBS,
                          'traffic': bsr_traffic, 'traffic_source': bsr_src}}
    print('# nearfield: %s' % json.dumps(res), flush=True)
    return res


def krylov_bench(n=978432, ms=(100, 500, 1000), reps=5):
    """One CGS2 orthogonalisation step (tct_cgs2: four or six launches) of
    the device GMRES at the large fault/topography solve's size (n free
    DOFs), for basis sizes m.  HBM-bound: the step reads the basis V[:m]
    in three passes at least (the first sums, the fused first update and
    second sums, the fused second update and norm), 24 m n bytes; the
    fused launches read their block's slice of V twice (update, then sums),
    so four passes, 32 m n bytes, are read unless the second read hits in
    cache.  Reported against 8 TB/s with the three-pass count."""
    import ctypes
    import torch
    from tectosaur_amd import _native as N
    out = {'workload': 'CGS2 step, n = %d (the large solve\'s free DOFs), FP64' % n}
    mmax = max(ms)
    V = torch.empty((mmax, n), dtype=torch.float64, device='cuda').normal_()
    V /= torch.linalg.vector_norm(V, dim=1, keepdim=True)
    w0 = torch.empty(n, dtype=torch.float64, device='cuda').normal_()
    for m in ms:
        nd = ctypes.c_int64(0)
        N.call('tct_krylov_scratch', m, n, ctypes.byref(nd))
        scratch = torch.empty(nd.value, dtype=torch.float64, device='cuda')
        h = torch.empty(m + 1, dtype=torch.float64, device='cuda')
        w = w0.clone()

        def step():
            w.copy_(w0)
            N.call('tct_cgs2', m, n, n, N.ptr(V), N.ptr(w), N.ptr(h), N.ptr(scratch), nd.value,
                   None, None, N.stream_ptr())
        step()
        dt = stream_event_time(step, reps)
        # the timed step includes the reset of w (16 n bytes, under 1% of
        # the step's bytes for m >= 100)
        alg = 24.0 * m * n
        out['m%d' % m] = {'ms': dt * 1e3, 'GB_per_s_3pass': alg / dt / 1e9,
                          'frac_hbm_peak_3pass': alg / dt / 1e9 / PEAK_HBM_GBS,
                          'GB_per_s_4pass': 32.0 * m * n / dt / 1e9}
    del V
    torch.cuda.empty_cache()
    print('# krylov: %s' % json.dumps(out), flush=True)
    return out


def treecode_bench(pts, tris, v, settings=((3, 2.0), (4, 2.5))):
    """The treecode farfield (FMMFarfieldOp) against the direct N-body on the
    same workload, RH3 kernel (the served operator's farfield): time per
    matvec and relative L2 error against the direct FP64 matvec."""
    import torch
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    n = tris.shape[0]
    idx = np.arange(n)
    op = TriToTriDirectFarfieldOp(1, 'elasticRH3', PARAMS, pts, tris, np.float32, idx, idx,
                                  precisions=('f32', 'f64'))
    v_dev = torch.as_tensor(v, dtype=torch.float64, device='cuda')
    y64 = op.dot_dev(v_dev, prec='f64')
    nrm = torch.linalg.norm(y64).item()
    res = {'kernel': 'elasticRH3', 'direct_f64_ms': 1e3 * stream_event_time(
               lambda: op.dot_dev(v_dev, prec='f64'), 1),
           'direct_f32_ms': 1e3 * stream_event_time(lambda: op.dot_dev(v_dev, prec='f32'), 2)}
    for order, mac in settings:
        t0 = time.perf_counter()
        tr = op.attach_tree(order=order, mac=mac)
        t_build = time.perf_counter() - t0
        yt = op.dot_dev(v_dev, prec='tree')
        err = (torch.linalg.norm(yt - y64) / nrm).item()
        ms = 1e3 * stream_event_time(lambda: op.dot_dev(v_dev, prec='tree'), 3)
        res['tree_o%d_mac%g' % (order, mac)] = {
            'ms': ms, 'rel_l2_vs_f64_direct': err, 'speedup_vs_direct_f32': res['direct_f32_ms'] / ms,
            'host_setup_s': t_build, 'm2p_nodes': tr.stats['m2p'], 'p2p_tiles': tr.stats['p2p_tiles'],
            'segments': tr.n_seg}
    print('# treecode: %s' % json.dumps(res), flush=True)
    return res


def interior_bench(n_grid=250, nq_far=3, nq_near=8, nq_vertex=8, reps=5):
    """Interior evaluation (InteriorOp): displacement (elasticT3) of the
    2,048-triangle vertical fault (x in [-1000, 1000] m, z in [-2000, 0] m,
    33 x 33 points) on an n_grid x n_grid horizontal grid at z = -500 m over
    [-3000, 3000]^2 m that crosses the fault; slip N(0, 1) from
    default_rng(2).  FP32 farfield, FP64 pair blocks."""
    import torch
    from tectosaur_amd import mesh
    from tectosaur_amd.interior import InteriorOp
    pts, tris = mesh.make_rect(33, 33, [[-1000, 0, 0], [1000, 0, 0], [1000, 0, -2000],
                                        [-1000, 0, -2000]])
    xs = np.linspace(-3000, 3000, n_grid)
    X, Y = np.meshgrid(xs, xs)
    obs = np.stack([X.ravel(), Y.ravel(), np.full(X.size, -500.0)], axis=1)
    ns = np.tile([0.0, 0.0, 1.0], (obs.shape[0], 1))
    v = np.random.default_rng(2).standard_normal(tris.shape[0] * 9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    op = InteriorOp(obs, ns, (pts, tris), 'elasticT3', 4.0, nq_vertex, nq_far, nq_near, PARAMS,
                    np.float32)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    v_dev = torch.as_tensor(v, dtype=torch.float64, device='cuda')
    out = torch.empty(op.shape[0], dtype=torch.float64, device='cuda')
    op.dot_dev(v_dev, out)
    ms = 1e3 * stream_event_time(lambda: op.dot_dev(v_dev, out), reps)
    far_ms = 1e3 * stream_event_time(lambda: op.farfield.dot_dev(v_dev, out), reps)
    pe = op.pair_evals
    res = {'workload': 'elasticT3 of a 2,048-tri vertical fault at a %dx%d grid (%d points) '
                       'crossing it; far rule %d (%d pts), near rule %d, vertex rule %d, '
                       'threshold 4' % (n_grid, n_grid, obs.shape[0], nq_far,
                                        op.farfield.nq, nq_near, nq_vertex),
           'close_pairs': int(op.n_blocks), 'vertex_pairs': int(op.vertex_pairs.shape[0]),
           'setup_s': t_build, 'dot_ms': ms, 'farfield_ms': far_ms,
           'farfield_pair_evals_per_s': pe / (far_ms * 1e-3),
           'points_per_s': obs.shape[0] / (ms * 1e-3)}
    print('# interior: %s' % json.dumps(res), flush=True)
    return res


def pairfind_bench(pts, tris, thr=2.5, reps=3):
    """Near-pair finding and adjacency classification on the two-surface
    mesh (147,456 triangles, threshold 2.5): the device finder against the
    host C++/OpenMP finder; both must return the same pairs."""
    import torch
    from tectosaur_amd import find_near_adj as F
    ca, ra = F.get_tri_centroids_rs(pts, tris)
    t_dev = torch.as_tensor(np.ascontiguousarray(tris, dtype=np.int64), device='cuda')
    F.find_near_pairs_dev(ca, ra, ca, ra, thr)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        d = F.find_near_pairs_dev(ca, ra, ca, ra, thr)
        cls_d = F.split_adjacent_close_dev(d, t_dev, t_dev)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    th = []
    for _ in range(reps):
        t0 = time.perf_counter()
        h = F.find_near_pairs(ca, ra, ca, ra, thr)
        cls_h = F.split_adjacent_close(h, tris, tris)
        th.append(time.perf_counter() - t0)
    same = bool(np.array_equal(d.cpu().numpy(), h) and
                all(np.array_equal(a.cpu().numpy(), b) for a, b in zip(cls_d, cls_h)))
    res = {'workload': 'two-surface mesh, %d tris, threshold %g' % (tris.shape[0], thr),
           'pairs': int(h.shape[0]), 'classes': [int(x.shape[0]) for x in cls_h],
           'device_s': min(ts), 'host_s': min(th), 'host_threads': os.cpu_count(),
           'identical': same}
    print('# pair finding: %s' % json.dumps(res), flush=True)
    return res


def large_mesh():
    from tectosaur_amd import mesh
    return mesh.fault_topography_mesh()


def large_solve_bench(mode='tree', tol=1e-8):
    """One traction -> slip solve (RH3 near and far, continuity and
    free-edge constraints) on the 655,360-triangle fault-plus-topography
    mesh: FP32 treecode farfield in the Arnoldi cycles (the inner operator;
    the workload's own direct FP32 farfield costs ~3 s per matvec, so a
    solve with it as the inner operator would take ~17 min, reported as an
    estimate from the measured matvec and the solve's inner matvec count),
    FP64 direct residuals; traction (1, 0.5, -0.3) + N(0, 0.1) noise
    (default_rng(0)).  Reports set-up and solve times, iterations, residual
    evaluations, the FP32/FP64 direct and treecode matvec times and peak
    device memory."""
    import torch
    from tectosaur_amd.solver import solve_constrained
    from tectosaur_amd.traction_slip import (DEFAULT_TREE, TractionSlipProblem,
                                             prepare_problem)
    pts, tris = large_mesh()
    n = tris.shape[0]
    t = np.tile([1.0, 0.5, -0.3], 3 * n) + np.random.default_rng(0).normal(0, 0.1, 9 * n)
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    prep = prepare_problem(pts, tris, t, tree=dict(DEFAULT_TREE))
    t_prep = time.perf_counter() - t0
    print('# large solve: %d tris, host prep %.1f s' % (n, t_prep), flush=True)
    t0 = time.perf_counter()
    prob = TractionSlipProblem(None, None, None, farfield_precisions=('f32', 'f64'),
                               prepared=prep)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    print('# large solve: device build %.1f s' % t_build, flush=True)
    ff = prob.H.farfield
    v = torch.as_tensor(np.random.default_rng(1).standard_normal(9 * n), device='cuda')
    mv = {}
    for prec in ('tree', 'f32', 'f64'):
        ff.dot_dev(v, prec=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ff.dot_dev(v, prec=prec)
        torch.cuda.synchronize()
        mv[prec + '_farfield_ms'] = 1e3 * (time.perf_counter() - t0)
        print('# large solve: %s farfield matvec %.1f ms' % (prec, mv[prec + '_farfield_ms']),
              flush=True)
    t0 = time.perf_counter()

    def progress(rel):
        print('# large solve: residual %.3e at %.1f s' % (rel, time.perf_counter() - t0),
              flush=True)

    def step(it, est):
        # with the direct inner operator a cycle takes minutes: report its
        # iterations as they go
        if mode == 'mixed' and it % 16 == 0:
            print('# large solve: iteration %d, estimated residual %.2e at %.1f s' % (
                it, est, time.perf_counter() - t0), flush=True)
    u, res = solve_constrained(prob.H, prob.cm, prob.rhs, tol=tol, mode=mode, callback=progress,
                               step_callback=step)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    inner = {'tree': 'FP32 treecode (order %d, mac %g) + FP32 nearfield blocks' % (
                 DEFAULT_TREE['order'], DEFAULT_TREE['mac']),
             'mixed': 'FP32 direct N-body + FP64 nearfield blocks',
             'f64': 'FP64 direct N-body (no inner operator)'}[mode]
    out = {'workload': 'traction -> slip, %d-tri fault + topography mesh, RH3 near/far, '
                       'tol %g, mode %s' % (n, tol, mode),
           'inner_operator': inner,
           'residual_operator': 'FP64 direct farfield + FP64 nearfield',
           'direct_f32_inner_estimate_s': (mv['f32_farfield_ms'] * 1e-3 *
                                           (res.n_matvec.get('tree', 0) +
                                            res.n_matvec.get('f32', 0))),
           'n_tris': int(n), 'n_dofs_free': int(prob.n_dofs), 'host_prep_s': t_prep,
           'device_build_s': t_build, 'solve_s': t_solve, 'converged': bool(res.converged),
           'iterations': int(res.iterations), 'restarts': int(res.restarts),
           'rel_residual_fp64': float(res.rel_residual), 'matvecs': res.n_matvec,
           'pair_evals_per_direct_matvec': int(prob.pair_evals),
           'peak_device_GB': torch.cuda.max_memory_allocated() / 1e9}
    out.update(mv)
    print('# large solve: %s' % json.dumps(out), flush=True)
    return out


def cpu_baseline(pts, tris, v, rows, out64_rows, out32_rows, seconds=8.0):
    """The reported CPU baseline: the H-kernel point N-body of
    tectosaur/farfield_direct.cl:15-113 restated for a many-core CPU
    (oracle/cpu_farfield.c, -O3 -march=native -ffast-math, OpenMP over all
    host threads), FP32 and FP64, built here for this CPU and timed on a
    bounded sample of the workload's observation quadrature points against
    all 442,368 source quadrature points (the strengths the device N-body
    sums: J w phi-weighted slip).  The FP64 oracle (long double) checks the
    GPU FP32 and FP64 matvecs on ``rows``."""
    from oracle import cpu_baseline as CB
    from oracle import oracle as O
    from tectosaur_amd.quadrature import gauss2d_tri
    qx, qw = gauss2d_tri(1)
    # quadrature points, normals and strengths (matrix_free.cl:50-90 order)
    phi = np.stack([1 - qx[:, 0] - qx[:, 1], qx[:, 0], qx[:, 1]], axis=1)
    tv = pts[tris]
    un = np.cross(tv[:, 2] - tv[:, 0], tv[:, 2] - tv[:, 1])
    jac = np.linalg.norm(un, axis=1)
    nrm = un / jac[:, None]
    qp = np.einsum('qb,tbd->tqd', phi, tv).reshape((-1, 3))
    qn = np.repeat(nrm, qx.shape[0], axis=0)
    vin = v.reshape((-1, 3, 3))
    s = np.einsum('t,q,qb,tbd->tqd', jac, qw, phi, vin).reshape((-1, 3))
    lib = CB.build()
    eng = CB.HFarfieldCPU(lib, qp, qn, s, PARAMS)
    rng = np.random.default_rng(5)
    sample = rng.permutation(qp.shape[0])
    res = {}
    for dt in (np.float32, np.float64):
        rate, n_obs, secs = eng.rate(qp[sample], qn[sample], dt, seconds / 2)
        res[np.dtype(dt).name] = (rate, n_obs, secs)
    # the two precisions against each other on the common prefix
    m = min(res['float32'][1], res['float64'][1], 256)
    y32 = eng.run(qp[sample[:m]], qn[sample[:m]], np.float32).astype(np.float64)
    y64 = eng.run(qp[sample[:m]], qn[sample[:m]], np.float64)
    model, ncpu = CB.cpu_model()
    cores = int(os.environ.get('OMP_NUM_THREADS', ncpu or 1))
    t0 = time.perf_counter()
    ref = O.farfield_tris('elasticH3', qx, qw, pts, tris, tris, v, PARAMS, rows=rows)
    t_or = time.perf_counter() - t0
    rl2 = lambda a: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
    base = {'value': res['float32'][0], 'unit': 'pair-evals/s', 'cores': cores, 'kind': 'port',
            'sample': '%d observation quadrature points x all %d source points (H, FP32, '
                      '%.1f s); FP64: %d points, %.1f s' % (
                          res['float32'][1], qp.shape[0], res['float32'][2],
                          res['float64'][1], res['float64'][2]),
            'f64_pair_evals_per_s': res['float64'][0],
            'cpu_model': model, 'logical_cpus': ncpu,
            'build': '-O3 -march=native -ffast-math -fopenmp (oracle/cpu_farfield.c)',
            'f32_vs_f64_rel_l2': float(np.linalg.norm(y32 - y64) / np.linalg.norm(y64))}
    return base, {'f64_vs_oracle_rel_l2': rl2(out64_rows), 'f32_vs_oracle_rel_l2': rl2(out32_rows),
                  'oracle_rows': int(rows.shape[0]), 'oracle_s': t_or}


def serve_bench(n_req, max_tris, rate, seed=0, max_active=32, min_tris=512,
                modes=('tree', 'mixed', 'f64'), baseline=True):
    """Served traction -> slip solves: one Poisson stream of fault requests
    through the continuous-batching server and through the plain FP64
    one-at-a-time path (always run, the baseline); latency = finish -
    arrival (queueing, device setup and solve).  Every result's residual is
    the FP64 direct operator's.  Batched ``modes`` separate the gains:
    'f64' (batching alone: the plain FP64 operator, 32 slots), 'mixed'
    (+ FP32 direct farfield in the Arnoldi cycles with FP64 refinement) and
    'tree' (+ the FP32 treecode as the inner farfield)."""
    import torch
    from tectosaur_amd.serve import PREP_WORKERS, Server, summarize
    from tectosaur_amd.traction_slip import random_requests
    warm = random_requests(2, rate=0.0, seed=seed + 99, n_min=512, n_max=1024)
    for m in modes:
        base_mode, _, opt = m.partition('+')
        Server(mode=base_mode, max_active=2, batch_nbody=(opt == 'nbody')).run(warm,
                                                                              realtime=False)
    if baseline:
        Server(mode='f64', max_active=1, prep_workers=0, graphs=False).run(warm, realtime=False)
    reqs = random_requests(n_req, rate=rate, seed=seed, n_min=min_tris, n_max=max_tris)
    out = {'workload': '%d fault requests, Poisson %.2f/s, strike U(0,360) dip U(20,90) '
                       'length U(5,50) km aspect U(1,4), %d-%d tris log-uniform, RH3 near and'
                       ' far, tol 1e-8' % (n_req, rate, min_tris, max_tris)}
    # a mode with the suffix '+nbody' runs the FP32 direct-farfield N-bodies
    # of each tick's solves as one segment-table launch (Server batch_nbody)
    runs = [('batched_' + m.replace('+', '_'), m, max_active, PREP_WORKERS) for m in modes
            if m != 'seq']
    if baseline:
        runs.append(('sequential_f64', 'f64', 1, 0))
    for name, mode, ma, pw in runs:
        base_mode, _, opt = mode.partition('+')
        srv = Server(mode=base_mode, max_active=ma, prep_workers=pw, graphs=(ma > 1),
                     batch_nbody=(opt == 'nbody'))
        torch.cuda.synchronize()
        recs = srv.run(reqs, realtime=True)
        sm = summarize(recs, srv.makespan)
        sm['ticks'] = srv.ticks
        sm['loop_wait_s'] = srv.sync_wait_s   # ticking thread waiting for the device
        out[name] = sm
        print('# serve %s: %s' % (name, json.dumps(sm)), flush=True)
    if not baseline:
        return out
    f = out['sequential_f64']
    for name, _, _, _ in runs[:-1]:
        b = out[name]
        out[name]['speedup_solves_per_s'] = b['solves_per_s'] / f['solves_per_s']
        out[name]['p50_ratio'] = f['p50_s'] / b['p50_s']
        out[name]['p99_ratio'] = f['p99_s'] / b['p99_s']
    if 'batched_tree' in out:
        out['speedup_solves_per_s'] = out['batched_tree']['speedup_solves_per_s']
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-rows', type=int, default=512)
    ap.add_argument('--kernels', default='elasticU3,elasticT3,elasticA3,elasticH3')
    ap.add_argument('--serve-requests', type=int, default=256)
    ap.add_argument('--serve-modes', default='tree,mixed,f64',
                    help="batched modes to run beside the sequential FP64 path, from "
                         "'tree', 'mixed', 'f64'; '+nbody' (e.g. 'mixed+nbody') batches "
                         "each tick's FP32 direct-farfield N-bodies into one launch")
    ap.add_argument('--serve-max-tris', type=int, default=32768)
    ap.add_argument('--serve-rate', type=float, default=100.0)
    ap.add_argument('--no-serve', action='store_true')
    ap.add_argument('--no-serve-baseline', action='store_true',
                    help='skip the sequential FP64 served run (for traces of the batched path)')
    ap.add_argument('--no-nearfield', action='store_true')
    ap.add_argument('--no-treecode', action='store_true')
    ap.add_argument('--no-interior', action='store_true')
    ap.add_argument('--no-pairfind', action='store_true')
    ap.add_argument('--no-krylov', action='store_true')
    ap.add_argument('--large-solve', action='store_true',
                    help='only the single large solve (655,360 triangles; minutes)')
    ap.add_argument('--large-solve-mode', default='tree', choices=('tree', 'mixed'),
                    help="inner operator of the large solve: 'tree' (FP32 treecode) or "
                         "'mixed' (the direct FP32 farfield; ~15 min)")
    ap.add_argument('--valu-calibration', action='store_true',
                    help='time the VALU issue-rate calibration kernels first')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run a 1 GiB device copy first (scale for byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    torch.cuda.set_device(0)
    if args.large_solve:
        print(json.dumps({'large_solve': large_solve_bench(mode=args.large_solve_mode)}),
              flush=True)
        return
    calib = copy_calibration() if args.copy_calibration else None
    valu_cal = valu_calibration() if args.valu_calibration else None
    t_setup = time.perf_counter()
    pts, tris = two_surface_mesh()
    n = tris.shape[0]
    v = np.random.default_rng(0).standard_normal(n * 9)
    v_dev = torch.as_tensor(v, dtype=torch.float64, device='cuda')
    all_idx = np.arange(n)
    print('# mesh: %d tris, %d pts (setup %.1fs)' % (n, pts.shape[0],
                                                    time.perf_counter() - t_setup), flush=True)

    per_kernel = {}
    H_out = {}
    for K in args.kernels.split(','):
        op = TriToTriDirectFarfieldOp(1, K, PARAMS, pts, tris, np.float32, all_idx, all_idx,
                                      precisions=('f32', 'f64'))
        pe = op.pair_evals
        res = {}
        outs = {}
        for prec in ('f32', 'f64'):
            steps = args.steps if (K == 'elasticH3' and prec == 'f32') else 1
            dt, out, k_dt = time_matvec(op, v_dev, prec, steps, args.warmup if steps > 1 else 1)
            outs[prec] = out.cpu().numpy()
            tc = op.engine.tile_cfg[prec]
            res[prec] = {'pair_evals_per_s': pe / dt, 'ms_per_matvec': dt * 1e3,
                         'nbody_ms': k_dt * 1e3,
                         'tile': [tc.block, tc.obs_per_lane, tc.src_chunks, tc.flags]}
            print('# %s %s: %.3f ms/matvec, %.3e pair-evals/s' % (K, prec, dt * 1e3, pe / dt),
                  flush=True)
        res['f32_vs_f64_rel_l2'] = float(np.linalg.norm(outs['f32'] - outs['f64']) /
                                         np.linalg.norm(outs['f64']))
        per_kernel[K] = res
        if K == 'elasticH3':
            H_out = outs
            H_op = op
            H_dt = res['f32']['ms_per_matvec'] * 1e-3
            # the N-body's own duration inside the timed matvecs
            kt = res['f32']['nbody_ms'] * 1e-3
        else:
            del op
        torch.cuda.empty_cache()

    pe = H_op.pair_evals
    flops = 102.0 * pe
    achieved = flops / kt / 1e12
    H_tile = per_kernel['elasticH3']['f32']['tile']
    traffic, traffic_src = profiled_traffic('far_p2p_f32_kernel<H>', H_tile)
    line = {
        'metric': 'H-kernel pair-evals/s + served-solve p50/p99 vs plain FP64, at 1e-8 residual',
        'value': pe / H_dt,
        'unit': 'pair-evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': H_dt * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': 'farfield direct H matvec, 40 km free surface 256x256 squares'
                               ' + 60-deg fault 128x64 squares, 147,456 tris, 3-pt rule,'
                               ' 442,368 x 442,368 point pairs',
                   'n_tris': int(n), 'n_points': int(n * 3), 'pair_evals_per_step': int(pe),
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'valu', 'achieved': achieved, 'peak': PEAK_FP32_TFLOPS,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32_TFLOPS,
                     'traffic': traffic, 'traffic_source': traffic_src,
                     'algorithmic_bytes': int(n * 3 * (24 + 48 + 24 * H_tile[2])),
                     'kernel': 'far_p2p_f32_kernel<H>', 'kernel_ms': kt * 1e3,
                     'kernel_timing': 'HIP events around the N-body launch inside the %d '
                                      'timed matvecs' % args.steps,
                     'tile': H_tile, 'flops_per_pair': 102},
        'kernels': per_kernel,
    }
    if calib is not None:
        line['copy_calibration'] = calib
    if valu_cal is not None:
        line['valu_calibration'] = valu_cal
    if not args.no_treecode:
        line['treecode'] = treecode_bench(pts, tris, v)
    if not args.no_nearfield:
        line['nearfield'] = nearfield_bench()
    if not args.no_interior:
        line['interior'] = interior_bench()
    if not args.no_pairfind:
        line['pair_finding'] = pairfind_bench(pts, tris)
    if not args.no_krylov:
        line['krylov'] = krylov_bench()
    if not args.no_serve:
        line['served'] = serve_bench(args.serve_requests, args.serve_max_tris, args.serve_rate,
                                     modes=tuple(m for m in args.serve_modes.split(',') if m),
                                     baseline=not args.no_serve_baseline)
    if not args.no_cpu_baseline:
        rng = np.random.default_rng(1)
        rows = np.sort(rng.choice(n, size=min(args.cpu_rows, n), replace=False))
        sel = (rows[:, None] * 9 + np.arange(9)[None, :]).reshape(-1)
        base, errs = cpu_baseline(pts, tris, v, rows, H_out['f64'][sel], H_out['f32'][sel])
        line['cpu_baseline'] = base
        line['accuracy_vs_oracle'] = errs
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
