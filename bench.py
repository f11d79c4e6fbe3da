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
PARAMS = [3e10, 0.25]


def cfg2_mesh():
    from tectosaur_amd import mesh
    pts, tris, n_surf = mesh.fault_surface_mesh(257, 129, 65, surf_w=20000.0, fault_L=10000.0,
                                                dip_deg=60.0)
    return pts, tris


def time_matvec(op, v_dev, prec, steps, warmup):
    import torch
    out = torch.empty(op.shape[0], dtype=torch.float64, device=op.device)
    for _ in range(warmup):
        op.dot_dev(v_dev, out, prec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        op.dot_dev(v_dev, out, prec)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def kernel_event_time(op, prec, reps):
    """Average duration of the N-body kernel alone, HIP events on the stream
    the kernel is launched on (torch's current stream)."""
    import torch
    e = op.engine
    s = torch.cuda.current_stream()
    t = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        e.run(prec)
        b.record(s)
        b.synchronize()
        t.append(a.elapsed_time(b) * 1e-3)
    return float(np.mean(t))


def cpu_baseline(pts, tris, v, rows, out64_rows, out32_rows):
    """The CPU oracle (C restatement of farfield_tris_to_tris, FP64 with
    long-double accumulation, OpenMP) on a bounded sample of observation
    triangles.  Also returns the GPU FP64/FP32 error on those rows."""
    import subprocess
    from oracle import oracle as O
    from tectosaur_amd.quadrature import gauss2d_tri
    if not os.path.exists(os.path.join(ROOT, 'oracle', 'liboracle.so')):
        subprocess.run(['make', '-C', os.path.join(ROOT, 'oracle')], check=True)
    qx, qw = gauss2d_tri(1)
    cores = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    t0 = time.perf_counter()
    ref = O.farfield_tris('elasticH3', qx, qw, pts, tris, tris, v, PARAMS, rows=rows)
    dt = time.perf_counter() - t0
    pairs = rows.shape[0] * 3 * tris.shape[0] * 3
    rl2 = lambda a: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
    return {'value': pairs / dt, 'unit': 'pair-evals/s', 'cores': cores, 'kind': 'port',
            'sample': '%d observation triangles x all %d source triangles (H, FP64 oracle,'
                      ' %.1f s)' % (rows.shape[0], tris.shape[0], dt)}, \
        {'f64_vs_oracle_rel_l2': rl2(out64_rows), 'f32_vs_oracle_rel_l2': rl2(out32_rows)}


def serve_bench(n_req, max_tris, rate, seed=0, max_active=32, min_tris=512):
    """Served traction -> slip solves: one Poisson stream of fault requests
    through the continuous-batching mixed-precision server and through the
    plain FP64 one-at-a-time path; latency = finish - arrival (queueing,
    device setup and solve)."""
    import torch
    from tectosaur_amd.serve import Server, summarize
    from tectosaur_amd.traction_slip import random_requests
    warm = random_requests(2, rate=0.0, seed=seed + 99, n_min=512, n_max=1024)
    Server(mode='mixed', max_active=2).run(warm, realtime=False)
    Server(mode='f64', max_active=1, prep_workers=0).run(warm, realtime=False)
    reqs = random_requests(n_req, rate=rate, seed=seed, n_min=min_tris, n_max=max_tris)
    out = {'workload': '%d fault requests, Poisson %.2f/s, strike U(0,360) dip U(20,90) '
                       'length U(5,50) km aspect U(1,4), %d-%d tris log-uniform, RH3 near and'
                       ' far, tol 1e-8' % (n_req, rate, min_tris, max_tris)}
    for name, mode, ma, pw in (('batched_mixed', 'mixed', max_active, 4),
                               ('sequential_f64', 'f64', 1, 0)):
        srv = Server(mode=mode, max_active=ma, prep_workers=pw)
        torch.cuda.synchronize()
        recs = srv.run(reqs, realtime=True)
        sm = summarize(recs, srv.makespan)
        sm['ticks'] = srv.ticks
        out[name] = sm
        print('# serve %s: %s' % (name, json.dumps(sm)), flush=True)
    b, f = out['batched_mixed'], out['sequential_f64']
    out['speedup_solves_per_s'] = b['solves_per_s'] / f['solves_per_s']
    out['p50_ratio'] = f['p50_s'] / b['p50_s']
    out['p99_ratio'] = f['p99_s'] / b['p99_s']
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-rows', type=int, default=4096)
    ap.add_argument('--kernels', default='elasticU3,elasticT3,elasticA3,elasticH3')
    ap.add_argument('--serve-requests', type=int, default=48)
    ap.add_argument('--serve-max-tris', type=int, default=8192)
    ap.add_argument('--serve-rate', type=float, default=4.0)
    ap.add_argument('--no-serve', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from tectosaur_amd.sparse_farfield_op import TriToTriDirectFarfieldOp
    torch.cuda.set_device(0)
    t_setup = time.perf_counter()
    pts, tris = cfg2_mesh()
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
            dt, out = time_matvec(op, v_dev, prec, steps, args.warmup if steps > 1 else 1)
            outs[prec] = out.cpu().numpy()
            res[prec] = {'pair_evals_per_s': pe / dt, 'ms_per_matvec': dt * 1e3,
                         'tile': [op.engine.tile_cfg[prec].block,
                                  op.engine.tile_cfg[prec].obs_per_lane,
                                  op.engine.tile_cfg[prec].src_chunks]}
            print('# %s %s: %.3f ms/matvec, %.3e pair-evals/s' % (K, prec, dt * 1e3, pe / dt),
                  flush=True)
        res['f32_vs_f64_rel_l2'] = float(np.linalg.norm(outs['f32'] - outs['f64']) /
                                         np.linalg.norm(outs['f64']))
        per_kernel[K] = res
        if K == 'elasticH3':
            H_out = outs
            H_op = op
            H_dt = res['f32']['ms_per_matvec'] * 1e-3
            kt = kernel_event_time(op, 'f32', 3)
        else:
            del op
        torch.cuda.empty_cache()

    pe = H_op.pair_evals
    flops = 102.0 * pe
    achieved = flops / kt / 1e12
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
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32_TFLOPS, 'traffic': None,
                     'kernel': 'far_p2p_f32_kernel<H>', 'kernel_ms': kt * 1e3,
                     'flops_per_pair': 102},
        'kernels': per_kernel,
    }
    if not args.no_serve:
        line['served'] = serve_bench(args.serve_requests, args.serve_max_tris, args.serve_rate)
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
