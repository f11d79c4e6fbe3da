"""Continuous-batching server for traction -> slip solves on one GPU.

Requests (``traction_slip.FaultRequest``) arrive over time.  The loop
admits an arrived request as soon as a slot is free and its estimated
device footprint fits the free-HBM budget, builds its operators on the
device, and from then on keeps its whole Krylov state in HBM.  Every tick
advances a batch of active solves by one GMRES step: each solve enqueues
its device work (matvec as a HIP-graph replay, orthogonalisation) on its
solve stream, the small per-solve results (Hessenberg column or residual norm)
are concatenated on the device and cross to the host in one asynchronous
transfer, and the host-side Givens updates run.  The active solves are
split into two batches whose ticks alternate, so the device works on one
batch while the host consumes the other's results.  A converged (or
failed) solve leaves at once and its slot and memory are reused; requests
arriving meanwhile join the next batch issued.

Host-side set-up of arriving requests (mesh, pair classification, farfield
layout, mass and constraint matrices) runs on a pool of worker threads while
the device iterates the active solves; admission then only does the device
half (pair integration, packing, uploads), on a builder thread and a build
stream of its own.  Active solves are spread over a few solve streams, so
small solves' kernels overlap the large ones'.

Each solve's arithmetic does not depend on which other solves share its
ticks (its kernels read only its own buffers), so a result is the same alone
and in the mix.

``serve_sequential`` is the one-at-a-time reference path (plain FP64 by
default) the batched path is compared with.
"""
import collections
import concurrent.futures
import contextlib
import sys
import threading
import time

import numpy as np
import torch

from .solver import MODES, ConstrainedOperator, make_state
from .traction_slip import (DEFAULT_TREE, TractionSlipProblem, fault_mesh, prepare_problem,
                            request_traction)


class SolveRecord:
    def __init__(self, req):
        self.req = req
        self.rid = req.rid
        self.arrival = req.arrival
        self.admitted = None
        self.finished = None
        self.status = None
        self.iterations = 0
        self.restarts = 0
        self.rel_residual = None
        self.slip = None
        self.n_tris = None
        self.pair_evals = 0
        self.n_matvec = None
        self.setup_s = None
        self.prep_s = None

    @property
    def latency(self):
        return self.finished - self.arrival

    def as_dict(self):
        return dict(rid=self.rid, n_tris=self.n_tris, arrival=self.arrival,
                    admitted=self.admitted, finished=self.finished, latency=self.latency,
                    status=self.status, iterations=self.iterations, restarts=self.restarts,
                    rel_residual=self.rel_residual, setup_s=self.setup_s, prep_s=self.prep_s,
                    n_matvec=self.n_matvec)


def estimate_bytes(n_tris, restart, tree=False):
    """Device footprint of one solve: farfield layouts and partial sums,
    nearfield blocks (FP64 and the FP32 inner copy, ~90 pairs per triangle
    at the default threshold), constraint and mass matrices, Krylov basis,
    and the treecode's moments and segment partials when it has one."""
    n_pts = 3 * n_tris
    far = n_pts * (24 + 48 + 96 + 48 + 8 * 3 * 8 + 16)
    near = 90 * n_tris * (648 + 324 + 16)
    n_free = int(1.6 * n_tris) + 64
    krylov = (restart + 5) * n_free * 8
    tr = 80e6 + n_pts * 64 if tree else 0
    return int(far + near + krylov + 64 * n_tris * 8 + tr)


class _Active:
    def __init__(self, rec, prob, op, state, nbytes, stream=None, build_stream=None):
        self.rec = rec
        self.prob = prob
        self.op = op
        self.state = state
        self.nbytes = nbytes
        self.stream = stream  # the stream all of this solve's steps run on
        self.build_stream = build_stream  # the stream its buffers were allocated on
        self.busy = False     # a step is in flight


# A graph capture allocates from a private memory pool of torch's caching
# allocator.  When the graph is destroyed the pool's memory stays reserved
# and is never handed out again until the cache is emptied (or an
# allocation fails): ~0.3 GB per served solve, ~72 GB per 256-request
# stream, which left later streams of one process with too little free HBM
# to admit their solves.  A release empties the cache once this much
# reserved memory is unallocated.
CACHE_SLACK_BYTES = 32 << 30


def _release(a, capture_lock=None, cache_slack=None):
    """Drop a finished solve's device objects (and with them its HIP graphs)
    on the calling thread, never during a graph capture (``capture_lock``).
    Its buffers were allocated on its build stream and return to that
    stream's pool: the build stream first waits for whatever is still queued
    on the solve's stream, so a later build cannot reuse the memory under
    it.  With ``cache_slack`` (bytes), the caching allocator's cache is
    emptied, under the same lock, when more than that is reserved but
    unallocated (the memory pools of destroyed graphs among it)."""
    if a.build_stream is not None and a.stream is not None:
        a.build_stream.wait_stream(a.stream)
    with capture_lock or contextlib.nullcontext():
        a.op = None
        a.state = None
        a.prob = None
        if cache_slack is not None and (torch.cuda.memory_reserved() -
                                        torch.cuda.memory_allocated() > cache_slack):
            torch.cuda.empty_cache()


_PREP_LOCAL = threading.local()
# preparation workers and OpenMP threads per worker on the 16-core share of
# one GPU: on a 256-request stream of 512-32,768-triangle solves, 2 x 8 ran
# 43.0-45.9 solves/s (p50 1.7-1.9 s) against 40.1-41.9 (p50 2.2 s) with
# 4 x 4, 40-41 with 8 x 2 or 6 x 2, 42-43 with 3 x 4, 3 x 5 or 2 x 4, 34
# with 1 x 16: each request's preparation finishes sooner with more threads,
# and two workers keep the admissions fed
PREP_WORKERS = 2
PREP_OMP_THREADS = 8


def _prepare(req, orders, tree=None, omp_threads=None):
    if omp_threads and getattr(_PREP_LOCAL, 'omp', None) != omp_threads:
        from . import _native as N
        N.call('tct_set_host_threads', omp_threads)
        _PREP_LOCAL.omp = omp_threads
    t0 = time.perf_counter()
    pts, tris = fault_mesh(req)
    t = request_traction(req, tris.shape[0])
    prep = prepare_problem(pts, tris, t, orders=orders, tree=tree)
    prep['prep_s'] = time.perf_counter() - t0
    return prep


# Solves below this many triangles have their FP32 direct-farfield N-bodies
# batched per tick when the server's batch_nbody is set (the size below
# which the tree mode uses the direct farfield).
NBODY_BATCH_MAX_TRIS = 3000


def _build(prep, mode, tol, restart, maxiter, graphs=False, capture_stream=None,
           capture_lock=None, batch_nbody=False):
    precisions = ('f64',) if mode == 'f64' else ('f32', 'f64')
    prob = TractionSlipProblem(None, None, None, farfield_precisions=precisions, prepared=prep)
    op = ConstrainedOperator(prob.H, prob.cm)
    b = op.cmT.dot_dev(prob.rhs)
    if mode == 'tree' and prob.H.farfield.tree is None:
        # below the treecode's size threshold the inner operator is the FP32
        # direct farfield: refine as the mixed mode does (deeper cycles)
        mode = 'mixed'
    state = make_state(op, b, mode=mode, tol=tol, restart=restart, maxiter=maxiter)
    if graphs:
        precs = MODES[mode][0:2] + MODES[mode][3:4]
        # with batched N-bodies, the inner (FP32 direct) applies of small
        # solves are split graphs around the N-body; the other precisions,
        # and larger solves, keep whole graphs: a large solve's N-body fills
        # the device with its own tuned launch shape, and its stream
        # overlaps the others (on the 256-request stream, batching every
        # mixed-mode solve measured 11% fewer solves/s)
        split = batch_nbody and mode == 'mixed' and prob.n_tris < NBODY_BATCH_MAX_TRIS
        whole = tuple(p for p in precs if not (split and p == MODES[mode][0]))
        with capture_lock or contextlib.nullcontext():
            op.enable_graphs(whole, stream=capture_stream)
            if split:
                op.enable_split_graphs((MODES[mode][0],), stream=capture_stream)
        state.batch_split = split
    return prob, op, state


def _finalize(a, now, keep_slip):
    rec, st = a.rec, a.state
    rec.finished = now
    rec.status = st.status
    rec.iterations = st.it
    rec.restarts = st.restarts
    rec.rel_residual = st.rel
    rec.n_matvec = dict(a.op.n_matvec)
    if keep_slip:
        rec.slip = a.op.cm.dot_dev(st.x).cpu().numpy()


class Server:
    def __init__(self, mode='mixed', max_active=32, tol=1e-8, restart=1500, maxiter=6000,
                 hbm_fraction=0.8, orders=None, keep_slip=False, prep_workers=PREP_WORKERS, tree=None,
                 graphs=True, streams=4, switch_interval=1e-3, builders=1, batch_nbody=False,
                 cache_slack=CACHE_SLACK_BYTES):
        # device queries from worker threads start from cached values
        from . import _native as N
        from .autotune import device_arch
        device_arch(N.require_device())
        torch.cuda.device_count()
        self.mode = mode
        self.max_active = max_active
        self.tol = tol
        self.restart = restart
        self.maxiter = maxiter
        self.hbm_fraction = hbm_fraction
        # unallocated reserved memory above which a release empties the
        # allocator's cache (None: never)
        self.cache_slack = cache_slack
        self.orders = orders
        # mode 'tree': treecode parameters of the inner farfield
        self.tree = (tree or dict(DEFAULT_TREE)) if mode == 'tree' else None
        self.keep_slip = keep_slip
        self.prep_workers = prep_workers
        # the loop, the builder thread and the preparation threads share the
        # GIL: a short switch interval keeps the ticking loop responsive
        # (preparation in worker processes measured 2x slower: its results
        # are large and must be pickled back)
        self.switch_interval = switch_interval
        # capture each solve's operator applications into HIP graphs
        self.graphs = graphs
        # the FP32 direct N-bodies of all solves in a tick's batch that use
        # the direct inner farfield run as one segment-table launch
        # (farfield.run_batched) between each solve's split graphs; those
        # solves share one stream, the group stream
        self.batch_nbody = bool(batch_nbody and graphs)
        self.group_stream = torch.cuda.Stream() if self.batch_nbody else None
        self.nbody_launches = 0   # batched N-body launches issued
        # batched servers build new solves on a stream of their own
        # (default priority) so admissions overlap the running solves
        # (``builders`` threads, one build stream each; builds upload through
        # pinned memory and never wait for the device, so one thread keeps
        # up with the 256-request stream, and more measured no faster)
        n_b = builders if max_active > 1 else 0
        self.build_streams = [torch.cuda.Stream() for _ in range(n_b)]
        self._builder_local = threading.local()
        self._builder_count = 0
        self._builder_lock = threading.Lock()
        # batched servers spread the active solves over a few streams
        # (default priority): one solve's kernels stay in order on its
        # stream while small solves' kernels overlap the large ones'
        n_st = streams if max_active > 1 else 0
        self.streams = [torch.cuda.Stream() for _ in range(n_st)]
        # every graph capture of this server runs on this one stream (no
        # further pooled streams are taken while solves run), one capture
        # at a time
        self.capture_stream = torch.cuda.Stream() if graphs else None
        self.capture_lock = threading.Lock()
        self._next_stream = 0
        self.ticks = 0
        self.max_concurrent = 0   # most solves active in one tick
        self.sync_wait_s = 0.0    # loop time spent waiting for tick results

    def _init_builder(self):
        """Builder thread initialiser: the thread's own build stream."""
        with self._builder_lock:
            i = self._builder_count
            self._builder_count += 1
        self._builder_local.stream = self.build_streams[i % len(self.build_streams)]

    def _build_on_stream(self, prep):
        """Device build of one solve on the calling builder thread's build
        stream; returns its objects, an event the solve's stream waits on,
        and the build stream."""
        st = self._builder_local.stream
        with torch.cuda.stream(st):
            prob, op, state = _build(prep, self.mode, self.tol, self.restart, self.maxiter,
                                     self.graphs, self.capture_stream, self.capture_lock,
                                     self.batch_nbody)
            ev = torch.cuda.Event()
            ev.record()
        return prob, op, state, ev, st

    def _adopt(self, active, rec, prob, op, state, need, clock, ev=None, build_stream=None):
        rec.n_tris = prob.n_tris
        rec.pair_evals = prob.pair_evals
        rec.setup_s = clock() - rec.admitted
        stream = None
        if state.batch_split:
            stream = self.group_stream
        elif self.streams:
            stream = self.streams[self._next_stream % len(self.streams)]
            self._next_stream += 1
        if ev is not None:
            (stream or torch.cuda.current_stream()).wait_event(ev)
        state.stream = stream
        active.append(_Active(rec, prob, op, state, need, stream, build_stream))

    def _adopt_built(self, active, entry, clock):
        """Adopt a build from the builder queue, entry = (future, record,
        bytes).  Its objects are unpacked here so that afterwards only the
        active solve refers to them, and its release (on a builder thread,
        under the capture lock) is what destroys them and their graphs.
        Names bound in the serving loop's own frame would keep the last
        adopted solve alive past its release; rebinding them at the next
        adoption would then destroy its graphs on the ticking thread without
        the lock, possibly while a builder captures (a HIP error that aborts
        the process)."""
        fut, rec, need = entry
        prob, op, state, ev, bst = fut.result()
        self._adopt(active, rec, prob, op, state, need, clock, ev, bst)

    def _issue(self, batch):
        """Enqueue one step of every solve in ``batch`` (each on its own
        stream); then, in one native call, the main stream joins those
        streams, gathers the small per-solve results into one device buffer
        and copies it to pinned host memory without blocking.  Returns
        (batch, sizes, pinned view, event, outputs); the outputs are kept
        until consumed, after the event, so nothing they occupy is reused
        while the gather may still read it."""
        from . import _native as N
        from .farfield import run_batched
        outs = [None] * len(batch)
        used = {}
        group = []
        for i, a in enumerate(batch):
            a.busy = True
            if a.stream is not None:
                used[id(a.stream)] = a.stream
            if a.state.needs_nbody():
                group.append(i)
            else:
                outs[i] = a.state.issue()
        if group:
            # on the group stream: every member's pre graph, one N-body
            # launch for all of them, every member's post graph + CGS2
            with torch.cuda.stream(self.group_stream):
                for i in group:
                    batch[i].state.issue_pre()
                run_batched([batch[i].op.nbody_engine for i in group])
                self.nbody_launches += 1
                for i in group:
                    outs[i] = batch[i].state.issue_post()
        k = self._slot
        self._slot ^= 1
        src, ln = self._seg_src[k], self._seg_len[k]
        n = len(outs)
        if n > src.shape[0]:
            raise RuntimeError('batch larger than the gather table')
        for i, o in enumerate(outs):
            src[i] = o.data_ptr()
            ln[i] = o.numel()
        total = int(ln[:n].sum())
        if total > self._pinned[k].shape[0]:
            raise RuntimeError('batch results exceed the pinned buffer')
        sts = list(used.values())
        for i, st in enumerate(sts):
            self._join_st[i] = st.cuda_stream
            self._join_ev[i] = self._join_events[id(st)].cuda_event
        ev = self._events[k]
        N.call('tct_tick_gather', len(sts), self._join_st.ctypes.data, self._join_ev.ctypes.data,
               n, src.ctypes.data, ln.ctypes.data, self._staging[k].data_ptr(),
               self._staging[k].shape[0], self._pinned[k].data_ptr(),
               torch.cuda.current_stream().cuda_stream, ev.cuda_event)
        return batch, ln[:n].tolist(), self._pinned[k][:total], ev, outs

    def run(self, requests, realtime=True, log=None):
        """Serve ``requests`` (sorted by arrival).  With realtime=True a
        request is admitted no earlier than its arrival time after start;
        otherwise all are present at t = 0.  Admission is first-come
        first-served.  Returns SolveRecords."""
        pending = sorted(requests, key=lambda r: r.arrival)
        recs = {}
        active = []
        used = 0
        # the admission budget is a fraction of the free HBM: memory the
        # allocator holds cached (earlier runs' graph pools among it) is
        # released first so that it counts as free
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        free0, _ = torch.cuda.mem_get_info()
        budget = self.hbm_fraction * free0
        pool = (concurrent.futures.ThreadPoolExecutor(self.prep_workers)
                if self.prep_workers > 0 else None)
        old_switch = sys.getswitchinterval()
        if self.switch_interval:
            sys.setswitchinterval(self.switch_interval)
        futures = {}
        max_ahead = self.max_active + 2 * max(self.prep_workers, 1)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        qi = 0    # next request to admit
        si = 0    # next request to hand to the preparation pool
        inflight = collections.deque()
        building = collections.deque()
        # per batch at most max_active solves x (restart + 2) values; two of
        # everything (two batches in flight), allocated up front: no host
        # allocation (an unsafe call during another thread's graph capture)
        # happens in the loop
        cap = self.max_active * (self.restart + 2)
        self._pinned = [torch.empty(cap, dtype=torch.float64, pin_memory=True) for _ in range(2)]
        self._staging = [torch.empty(cap, dtype=torch.float64, device='cuda') for _ in range(2)]
        self._seg_src = [np.zeros(self.max_active, dtype=np.uint64) for _ in range(2)]
        self._seg_len = [np.zeros(self.max_active, dtype=np.int32) for _ in range(2)]
        self._events = [torch.cuda.Event() for _ in range(2)]
        # one join event per solve stream; each event is recorded once here
        # so its handle exists before the native gather re-records it
        all_streams = self.streams + ([self.group_stream] if self.group_stream else [])
        self._join_events = {id(st): torch.cuda.Event() for st in all_streams}
        self._join_st = np.zeros(max(1, len(all_streams)), dtype=np.uint64)
        self._join_ev = np.zeros(max(1, len(all_streams)), dtype=np.uint64)
        for st in all_streams:
            self._join_events[id(st)].record(st)
        for e in self._events:
            e.record()
        self._slot = 0
        builder = (concurrent.futures.ThreadPoolExecutor(len(self.build_streams),
                                                          initializer=self._init_builder)
                   if self.build_streams else None)
        try:
            while qi < len(pending) or active or building:
                now = clock()
                # host preparation of arrived requests, bounded look-ahead
                while pool is not None and si < len(pending) and si < qi + max_ahead:
                    if realtime and pending[si].arrival > now:
                        break
                    futures[pending[si].rid] = pool.submit(_prepare, pending[si], self.orders,
                                                           self.tree, PREP_OMP_THREADS)
                    si += 1
                # admission (FIFO) of prepared requests
                while (qi < len(pending) and
                       len(active) + len(building) < self.max_active):
                    req = pending[qi]
                    if realtime and req.arrival > now:
                        break
                    need = estimate_bytes(req.n_tris, self.restart,
                                          self.tree is not None and
                                          req.n_tris >= self.tree.get('min_tris', 0))
                    if (active or building) and used + need > budget:
                        break
                    if pool is not None:
                        fut = futures.get(req.rid)
                        if fut is None or ((active or building) and not fut.done()):
                            break
                        prep = fut.result()
                        del futures[req.rid]
                    else:
                        prep = None
                    rec = SolveRecord(req)
                    if not realtime:
                        rec.arrival = 0.0
                    rec.admitted = clock()
                    if prep is None:
                        prep = _prepare(req, self.orders, self.tree)
                    rec.prep_s = prep['prep_s']
                    recs[req.rid] = rec
                    used += need
                    qi += 1
                    if builder is not None:
                        # the device build runs on the builder thread and
                        # the build stream while this thread keeps ticking
                        building.append((builder.submit(self._build_on_stream, prep), rec,
                                         need))
                    else:
                        prob, op, state = _build(prep, self.mode, self.tol, self.restart,
                                                 self.maxiter, self.graphs, self.capture_stream,
                                                 batch_nbody=self.batch_nbody)
                        ev = None
                        if self.streams:
                            ev = torch.cuda.Event()
                            ev.record()
                        self._adopt(active, rec, prob, op, state, need, clock, ev)
                    now = clock()
                # adopt finished builds (block only when nothing else can run)
                while building and (building[0][0].done() or not active):
                    self._adopt_built(active, building.popleft(), clock)
                if not active:
                    if qi < len(pending):
                        time.sleep(max(0.0, min(0.002, pending[qi].arrival - clock())))
                    continue
                # Ticks are pipelined over (up to) two batches of solves: a
                # batch's step is issued, its small results head to pinned
                # host memory asynchronously, and while the host consumes one
                # batch (Givens updates, convergence) the device runs the
                # other batch's queued step.
                ready = [a for a in active if not a.busy]
                if ready and len(inflight) < 2:
                    if not inflight and len(ready) > 1:
                        half = (len(ready) + 1) // 2
                        inflight.append(self._issue(ready[:half]))
                        inflight.append(self._issue(ready[half:]))
                    else:
                        inflight.append(self._issue(ready))
                    continue
                self.max_concurrent = max(self.max_concurrent, len(active))
                members, sizes, pinned, ev, _outs = inflight.popleft()
                tw = time.perf_counter()
                ev.synchronize()
                self.sync_wait_s += time.perf_counter() - tw
                host = pinned.numpy()
                self.ticks += 1
                off = 0
                now = clock()
                finished = set()
                for a, sz in zip(members, sizes):
                    # device work the host step enqueues (basis vector,
                    # solution update) goes to the solve's own stream
                    done = a.state.consume(host[off:off + sz])
                    if done:
                        with torch.cuda.stream(a.stream or torch.cuda.current_stream()):
                            _finalize(a, now, self.keep_slip)
                    off += sz
                    a.busy = False
                    if done:
                        used -= a.nbytes
                        finished.add(id(a))
                        if log:
                            log(a.rec)
                        if builder is not None:
                            # the solve's graphs are destroyed on a builder
                            # thread, never while a capture is under way
                            builder.submit(_release, a, self.capture_lock,
                                           self.cache_slack)
                if finished:
                    active = [a for a in active if id(a) not in finished]
        finally:
            sys.setswitchinterval(old_switch)
            if pool is not None:
                pool.shutdown(wait=True, cancel_futures=True)
            if builder is not None:
                builder.shutdown(wait=True, cancel_futures=True)
        torch.cuda.synchronize()
        self.makespan = clock()
        return [recs[r.rid] for r in pending]


def serve_sequential(requests, mode='f64', realtime=True, **kw):
    """The one-at-a-time path: each request prepared, built and solved alone,
    in arrival order, with no overlap and no graph capture."""
    kw.setdefault('graphs', False)
    return Server(mode=mode, max_active=1, prep_workers=0, **kw).run(requests, realtime=realtime)


def summarize(recs, makespan):
    lat = np.array([r.latency for r in recs])
    ok = [r for r in recs if r.status == 'converged']
    return dict(n=len(recs), converged=len(ok), solves_per_s=len(recs) / makespan,
                makespan_s=makespan, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()),
                max_rel_residual=float(max(r.rel_residual for r in recs)),
                iterations_mean=float(np.mean([r.iterations for r in recs])))
