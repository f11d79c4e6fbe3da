"""Device-resident GMRES for constrained BEM systems.

The reference solves cm^T A cm x = cm^T (b - A c_rhs) with
scipy.sparse.linalg.gmres on the host, one host<->device round trip per
matvec (tectosaur/simple_solver.py:14-34, examples/solve.py:25-61).  Here
the Krylov vectors, the operator and the orthogonalisation stay in HBM:

  * ``ConstrainedOperator`` applies cm^T (A (cm x)) with device CSR matvecs
    around the summed farfield + nearfield matvec;
  * ``GmresState`` is restarted GMRES(m) with classical Gram-Schmidt
    repeated twice (CGS2) in FP64 on the device (the HIP kernels of
    krylov.hip for device vectors); only
    the Hessenberg column (m + 2 doubles) crosses to the host per iteration
    for the Givens update.  It is a step machine (``issue`` enqueues device
    work and returns a small device tensor, ``consume`` takes its host
    value), so a scheduler can advance many solves with one host sync;
  * mixed precision: the Arnoldi matvecs use the FP32 farfield and every
    restart recomputes the true residual b - A x with the FP64 farfield, so
    the restarts are iterative-refinement steps and the converged residual
    is the FP64 operator's.  A cycle stops once it has reduced the current
    residual by ``inner_reduction`` (the FP32 operator's accuracy bounds what
    one cycle can usefully gain).

Tolerances are relative to ||rhs|| of the constrained system, as scipy's.
"""
import contextlib
import ctypes
import math
import time

import numpy as np
import torch

from .mass_op import DeviceCSR


class ConstrainedOperator:
    """x -> cm^T A (cm x) on the device; A has dot_dev(v, out, prec)."""

    def __init__(self, A, cm):
        self.A = A
        self.cm = DeviceCSR(cm)
        self.cmT = DeviceCSR(cm.T.tocsr())
        self.device = self.cm.device
        self.shape = (cm.shape[1], cm.shape[1])
        self._u = torch.empty(cm.shape[0], dtype=torch.float64, device=self.device)
        self._y = torch.empty(cm.shape[0], dtype=torch.float64, device=self.device)
        self.n_matvec = {'f32': 0, 'f64': 0, 'tree': 0}
        self._graphs = {}
        self._gx = self._gy = None

    def _apply_raw(self, x, out, prec):
        self.cm.dot_dev(x, self._u)
        self.A.dot_dev(self._u, self._y, prec)
        self.cmT.dot_dev(self._y, out)

    def enable_graphs(self, precs, stream=None):
        """Capture the whole apply (CSR, farfield passes, BSR, CSR) for each
        precision in ``precs`` into a HIP graph replayed by ``apply``: one
        graph launch and two vector copies per application instead of a
        dozen kernel launches.  The captured kernels write every output in
        full (no memsets) and use only buffers the operator owns.

        Capture runs on a side stream ordered after the current one, without
        a device-wide synchronisation, so a server can build a new solve's
        graphs while other solves' work is in flight.  A server passes its
        one capture stream: torch hands out pooled streams round-robin, so a
        fresh stream per solve would in time alias a stream other solves
        replay on."""
        n = self.shape[0]
        if self._gx is None:
            # every graph of this operator reads and writes these two
            # buffers: allocated once, so later captures keep earlier
            # graphs valid
            self._gx = torch.zeros(n, dtype=torch.float64, device=self.device)
            self._gy = torch.empty(n, dtype=torch.float64, device=self.device)
        cur = torch.cuda.current_stream(self.device)
        side = stream if stream is not None else torch.cuda.Stream(self.device)
        for prec in precs:
            if prec is None or prec in self._graphs:
                continue
            if hasattr(self.A, 'prepare'):
                self.A.prepare(prec)                       # lazy buffers, no work
            else:
                self._apply_raw(self._gx, self._gy, prec)  # allocate them by running once
            side.wait_stream(cur)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(side):
                # thread-local: other host threads may keep launching work
                # on other streams while this one captures
                g.capture_begin(capture_error_mode='thread_local')
                try:
                    self._apply_raw(self._gx, self._gy, prec)
                finally:
                    g.capture_end()
            cur.wait_stream(side)
            self._graphs[prec] = g

    def enable_split_graphs(self, precs, stream=None):
        """Capture each apply of ``precs`` in two graphs around the farfield
        N-body, which a server then runs for many solves in one launch
        (farfield.run_batched): 'pre' = cm x, the farfield gather; 'post' =
        the farfield scatter, the nearfield product, cm^T.  The kernels and
        buffers are those of the whole apply, so pre + N-body + post gives
        its result bit for bit.  The farfield must be the direct N-body
        (TriToTriDirectFarfieldOp)."""
        ff = getattr(self.A, 'farfield', None)
        if ff is None or not hasattr(ff, '_gather') or not hasattr(ff, 'engine'):
            raise ValueError('split graphs need a direct farfield operator')
        n = self.shape[0]
        if self._gx is None:
            self._gx = torch.zeros(n, dtype=torch.float64, device=self.device)
            self._gy = torch.empty(n, dtype=torch.float64, device=self.device)
        if not hasattr(self, '_split'):
            self._split = {}
        cur = torch.cuda.current_stream(self.device)
        side = stream if stream is not None else torch.cuda.Stream(self.device)
        for prec in precs:
            if prec is None or prec in self._split or prec == 'tree':
                continue
            self.A.prepare(prec)
            pair = []
            for part in (self._split_pre, self._split_post):
                side.wait_stream(cur)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.stream(side):
                    g.capture_begin(capture_error_mode='thread_local')
                    try:
                        part(prec)
                    finally:
                        g.capture_end()
                cur.wait_stream(side)
                pair.append(g)
            self._split[prec] = tuple(pair)

    def _split_pre(self, prec):
        self.cm.dot_dev(self._gx, self._u)
        self.A.farfield._gather(self._u, prec)

    def _split_post(self, prec):
        ff = self.A.farfield
        ff._scatter(prec, self._y)
        self.A.nearfield.dot_dev(self._u, self._y, accumulate=True, prec=prec)
        self.cmT.dot_dev(self._y, self._gy)

    def has_split(self, prec):
        return prec in getattr(self, '_split', {})

    def split_pre_replay(self, prec):
        """The 'pre' graph of ``prec`` on the current stream."""
        self._split[prec][0].replay()

    def split_post_exec(self, prec):
        """The executable 'post' graph of ``prec`` (a hipGraphExec_t address)."""
        return self._split[prec][1].raw_cuda_graph_exec()

    @property
    def nbody_engine(self):
        return self.A.farfield.engine

    def has_graph(self, prec):
        return prec in self._graphs

    def graph_exec(self, prec):
        """The executable graph of ``prec`` (a hipGraphExec_t address)."""
        return self._graphs[prec].raw_cuda_graph_exec()

    def replay(self, prec):
        """One captured application: graph_input -> graph_output (the
        buffers ``gx``/``gy``), with no copies around it."""
        self._graphs[prec].replay()
        self.n_matvec[prec] += 1

    @property
    def gx(self):
        return self._gx

    @property
    def gy(self):
        return self._gy

    def apply(self, x, out=None, prec='f64'):
        if out is None:
            out = torch.empty(self.shape[0], dtype=torch.float64, device=self.device)
        g = self._graphs.get(prec)
        if g is not None:
            self._gx.copy_(x)
            g.replay()
            out.copy_(self._gy)
        else:
            self._apply_raw(x, out, prec)
        self.n_matvec[prec] += 1
        return out


def _givens(a, b):
    if b == 0.0:
        return 1.0, 0.0
    r = math.hypot(a, b)
    return a / r, b / r


class GmresResult:
    def __init__(self, x, converged, iterations, restarts, residuals, rel_residual, seconds,
                 status):
        self.x = x
        self.converged = converged
        self.iterations = iterations
        self.restarts = restarts
        self.residuals = residuals
        self.rel_residual = rel_residual
        self.seconds = seconds
        self.status = status  # 'converged', 'maxiter', 'nonfinite', 'breakdown'
        self.n_matvec = None

    @property
    def info(self):
        return 0 if self.converged else max(self.iterations, 1)


class GmresState:
    """Restarted GMRES as a step machine.

    apply(x, out, prec) computes the operator on device tensors.  Drive it
    with ``while not s.done: s.consume(s.issue().cpu())`` (what ``run``
    does), or interleave many states with one transfer per tick."""

    def __init__(self, apply, b, tol=1e-8, restart=200, maxiter=2000, inner_prec='f64',
                 outer_prec='f64', inner_reduction=None, x0=None, callback=None,
                 coarse_prec=None, coarse_above=3e-5, update_prec=None, update_below=1e-3,
                 operator=None, step_callback=None):
        self.apply = apply
        self.b = b
        dev = b.device
        n = b.shape[0]
        self.n = n
        self.m = max(1, min(restart, n))
        self.tol = tol
        self.maxiter = maxiter
        self.inner_prec = inner_prec
        self.outer_prec = outer_prec
        self.inner_reduction = inner_reduction
        # residuals expected above coarse_above * ||b|| may be computed in
        # coarse_prec (a cheaper operator accurate to well below that level);
        # convergence is only ever declared from an outer_prec residual
        self.coarse_prec = coarse_prec
        self.coarse_above = coarse_above
        # once an outer_prec residual is below update_below * ||b||, later
        # residuals may be updated, r <- r - A d with d the cycle's
        # correction, in update_prec: the update's error is relative to
        # |A d| ~ |r|, so a cheap operator keeps the residual accurate to far
        # below the target; a residual that meets the target this way is
        # confirmed with an outer_prec one before convergence is declared
        self.update_prec = update_prec
        self.update_below = update_below
        self._anchored = False    # the residual chain starts from an outer_prec one
        self._confirm = False     # the next residual must be outer_prec
        self._expect = float('inf')
        self._res_prec = outer_prec
        self.callback = callback
        # called with (iterations, estimated relative residual) after every
        # Arnoldi step
        self.step_callback = step_callback
        self.x = torch.zeros(n, dtype=torch.float64, device=dev) if x0 is None else x0.clone()
        self._fresh = x0 is None
        self.V = torch.empty((self.m + 1, n), dtype=torch.float64, device=dev)
        self.w = torch.empty(n, dtype=torch.float64, device=dev)
        self.r = torch.empty(n, dtype=torch.float64, device=dev)
        self.d = torch.zeros(n, dtype=torch.float64, device=dev) if update_prec else None
        # device solves orthogonalise with the workspace-free HIP kernels of
        # krylov.hip (fixed reduction order, nothing shared between solves
        # on different streams); host tensors use torch
        self._hip = dev.type == 'cuda'
        # with the operator's inner-precision graph, an Arnoldi step is one
        # replay reading the graph input buffer and writing the graph output,
        # which is orthogonalised in place; the next basis vector goes
        # straight into V and the graph input (no copies, no host step)
        self._op = operator
        self._fast = False
        # the stream all of this solve's device work goes to (a server sets
        # it); None: the current stream at each call
        self.stream = None
        # a server that runs the farfield N-body of many solves in one launch
        # sets this: Arnoldi steps then go pre graph -> (the server's batched
        # N-body) -> post graph + CGS2 (issue_pre / issue_post)
        self.batch_split = False
        self._split = False
        if self._hip:
            from . import _native as N
            nd = ctypes.c_int64(0)
            N.call('tct_krylov_scratch', self.m + 1, n, ctypes.byref(nd))
            self._scratch = torch.empty(nd.value, dtype=torch.float64, device=dev)
            self._h = torch.empty(self.m + 2, dtype=torch.float64, device=dev)
            # the cycle's least-squares solution goes up through pinned
            # memory without a stream synchronisation; the pinned buffer is
            # rewritten only at the next cycle's end, after later work on the
            # same stream has been waited for
            self._ypin = torch.empty(self.m, dtype=torch.float64, pin_memory=True)
            self._ydev = torch.empty(self.m, dtype=torch.float64, device=dev)
        self.it = 0
        self.restarts = 0
        self.residuals = []
        self.rel = float('inf')
        self.done = False
        self.status = None
        self.phase = 'norm_b'
        self.t0 = time.perf_counter()

    @property
    def device_bytes(self):
        return (self.m + 5) * self.n * 8

    # -- device side --------------------------------------------------------
    def _ctx(self):
        return (torch.cuda.stream(self.stream) if self.stream is not None
                else contextlib.nullcontext())

    def needs_nbody(self):
        """True when the next step is an Arnoldi step on split graphs: the
        caller runs issue_pre, the farfield N-body (batched with other
        solves') and then issue_post in place of issue."""
        return self.phase == 'arnoldi' and self._split and not self.done

    def issue_pre(self):
        """The part of a split Arnoldi step before the N-body."""
        with self._ctx():
            self._op.split_pre_replay(self.inner_prec)

    def issue_post(self):
        """The rest of a split Arnoldi step after the N-body: the post graph
        and CGS2 (one native call); returns what ``issue`` returns."""
        return self._arnoldi_native()

    def issue(self):
        """Enqueue the next step's device work (on ``stream`` when set, else
        the current stream); return the device tensor whose host value
        ``consume`` needs."""
        if self.phase == 'arnoldi' and self._split:
            with self._ctx():
                self._op.split_pre_replay(self.inner_prec)
                from .farfield import run_batched
                run_batched([self._op.nbody_engine])
            return self._arnoldi_native()
        if self.phase == 'arnoldi' and self._fast:
            return self._arnoldi_native()
        with self._ctx():
            return self._issue()

    def _arnoldi_native(self):
        # graph launch + CGS2 in one native call; the next basis vector goes
        # straight into V[j+1] and the graph input
        from . import _native as N
        j = self.j
        sp = self.stream.cuda_stream if self.stream is not None else N.stream_ptr()
        N.call('tct_arnoldi_step', self._gexec, j + 1, self.n, self.n, self._pV, self._pgy,
               self._ph, self._psc, self._nsc, self._pV + (j + 1) * self._row_bytes,
               self._pgx, sp)
        self._op.n_matvec[self.inner_prec] += 1
        return self._h[:j + 2]

    def _issue(self):
        if self.phase == 'norm_b':
            return torch.linalg.vector_norm(self.b).reshape(1)
        if self.phase == 'residual':
            self._res_kind = 'full'
            if self._fresh and self.restarts == 0:
                self.r.copy_(self.b)
                self._res_prec = self.outer_prec
            elif (self.update_prec is not None and self._anchored and not self._confirm and
                  self.rel <= self.update_below):
                # r <- r - A d (w is free between cycles)
                self._res_prec = self.update_prec
                self._res_kind = 'update'
                self.apply(self.d, self.w, self.update_prec)
                self.r.sub_(self.w)
            else:
                coarse = (self.coarse_prec is not None and not self._confirm and
                          self._expect > self.coarse_above)
                self._res_prec = self.coarse_prec if coarse else self.outer_prec
                self.apply(self.x, self.r, self._res_prec)
                torch.sub(self.b, self.r, out=self.r)
            return torch.linalg.vector_norm(self.r).reshape(1)
        # arnoldi step j (without the operator's graph)
        j = self.j
        self.apply(self.V[j], self.w, self.inner_prec)
        if self._hip:
            from . import _native as N
            N.call('tct_cgs2', j + 1, self.n, self.n, N.ptr(self.V), N.ptr(self.w),
                   N.ptr(self._h), N.ptr(self._scratch), self._scratch.shape[0], None, None,
                   N.stream_ptr())
            return self._h[:j + 2]
        Vj = self.V[:j + 1]
        h1 = torch.mv(Vj, self.w)
        self.w.sub_(torch.mv(Vj.t(), h1))
        h2 = torch.mv(Vj, self.w)
        self.w.sub_(torch.mv(Vj.t(), h2))
        return torch.cat([h1 + h2, torch.linalg.vector_norm(self.w).reshape(1)])

    # -- host side ----------------------------------------------------------
    def consume(self, vals):
        """Take the host values of the last issued step.  Device work this
        enqueues (basis vector, solution update) runs on ``stream`` when
        set; a mid-cycle Arnoldi step of the graph path enqueues none."""
        vals = np.asarray(vals, dtype=np.float64).reshape(-1)
        if self.phase == 'arnoldi' and self._fast:
            return self._consume_arnoldi(vals)
        with self._ctx():
            return self._consume(vals)

    def _consume(self, vals):
        if self.phase == 'norm_b':
            self.bnorm = float(vals[0])
            if self.bnorm == 0.0 or not math.isfinite(self.bnorm):
                self.x.zero_()
                ok = self.bnorm == 0.0
                self.rel = 0.0 if ok else float('nan')
                self.residuals.append(self.rel)
                return self._finish('converged' if ok else 'nonfinite')
            self.target = self.tol * self.bnorm
            self.phase = 'residual'
            return False
        if self.phase == 'residual':
            beta = float(vals[0])
            self.rel = beta / self.bnorm
            self.residuals.append(self.rel)
            if self.callback is not None:
                self.callback(self.rel)
            if not math.isfinite(beta):
                return self._finish('nonfinite')
            exact = self._res_prec == self.outer_prec and self._res_kind == 'full'
            if exact:
                self._anchored = True
            elif self._res_kind == 'full':
                self._anchored = False
            self._confirm = False
            if beta <= self.target:
                if not exact:
                    # a coarse or updated residual never decides convergence
                    self._expect = 0.0
                    self._confirm = True
                    return False
                return self._finish('converged')
            if self.it >= self.maxiter:
                return self._finish('maxiter')
            self.restarts += 1
            self.V[0].copy_(self.r).div_(beta)
            self._split = (self._hip and self._op is not None and self.batch_split and
                           self._op.has_split(self.inner_prec))
            self._fast = self._split or (self._hip and self._op is not None and
                                         self._op.has_graph(self.inner_prec))
            if self._fast:
                op = self._op
                op.gx.copy_(self.V[0])
                # addresses of the native Arnoldi step (with split graphs,
                # the step's graph is the 'post' one)
                self._gexec = (op.split_post_exec(self.inner_prec) if self._split else
                               op.graph_exec(self.inner_prec))
                self._pV, self._pgx, self._pgy = (self.V.data_ptr(), op.gx.data_ptr(),
                                                  op.gy.data_ptr())
                self._ph, self._psc = self._h.data_ptr(), self._scratch.data_ptr()
                self._nsc = self._scratch.shape[0]
                self._row_bytes = self.V.stride(0) * self.V.element_size()
            # Givens state in Python floats (the per-step host work is a few
            # dozen scalar operations); R's columns kept as lists
            self.Rcols = []
            self.g = [beta]
            self.cs = []
            self.sn = []
            self.cycle_target = self.target
            if self.inner_reduction is not None:
                self.cycle_target = max(self.target * 0.5, self.inner_reduction * beta)
            self.j = 0
            self.phase = 'arnoldi'
            return False
        return self._consume_arnoldi(vals)

    def _consume_arnoldi(self, vals):
        j = self.j
        col = vals.tolist()
        hnext = col[j + 1]
        self.it += 1
        cs, sn = self.cs, self.sn
        for i in range(j):
            a, b = col[i], col[i + 1]
            col[i] = cs[i] * a + sn[i] * b
            col[i + 1] = -sn[i] * a + cs[i] * b
        c, sv = _givens(col[j], col[j + 1])
        cs.append(c)
        sn.append(sv)
        col[j] = c * col[j] + sv * col[j + 1]
        self.Rcols.append(col[:j + 1])
        gj = self.g[j]
        self.g[j] = c * gj
        self.g.append(-sv * gj)
        est = abs(self.g[j + 1])
        if self.step_callback is not None:
            self.step_callback(self.it, est / self.bnorm)
        k = j + 1
        end = (not math.isfinite(est) or est <= self.cycle_target or hnext == 0.0 or
               k >= self.m or self.it >= self.maxiter)
        if not end:
            if not self._fast:
                self.V[k].copy_(self.w).div_(hnext)
            self.j = k
            return False
        # close the cycle: x += V_k y
        self._expect = est / self.bnorm if math.isfinite(est) else float('inf')
        with self._ctx():
            self._close_cycle(est, k)
        self.phase = 'residual'
        return False

    def _close_cycle(self, est, k):
        if math.isfinite(est):
            R = np.zeros((k, k))
            for jj, cj in enumerate(self.Rcols):
                R[:jj + 1, jj] = cj
            y = np.linalg.solve(R, np.asarray(self.g[:k]))
            if np.all(np.isfinite(y)):
                if self._hip:
                    from . import _native as N
                    self._ypin[:k].copy_(torch.from_numpy(y))
                    y_t = self._ydev[:k]
                    y_t.copy_(self._ypin[:k], non_blocking=True)
                    tgt, beta = (self.d, 0.0) if self.d is not None else (self.x, 1.0)
                    N.call('tct_krylov_combine', k, self.n, self.n, N.ptr(self.V), N.ptr(y_t),
                           1.0, beta, N.ptr(tgt), N.stream_ptr())
                    if self.d is not None:
                        self.x.add_(self.d)
                else:
                    y_t = torch.as_tensor(y, dtype=torch.float64, device=self.x.device)
                    if self.d is not None:
                        torch.mv(self.V[:k].t(), y_t, out=self.d)
                        self.x.add_(self.d)
                    else:
                        self.x.add_(torch.mv(self.V[:k].t(), y_t))
            elif self.d is not None:
                self.d.zero_()
        elif self.d is not None:
            self.d.zero_()

    def _finish(self, status):
        self.done = True
        self.status = status
        return True

    def result(self):
        return GmresResult(self.x, self.status == 'converged', self.it, self.restarts,
                           self.residuals, self.rel, time.perf_counter() - self.t0, self.status)

    def run(self):
        while not self.done:
            self.consume(self.issue().cpu().numpy())
        return self.result()


def gmres(apply, b, x0=None, tol=1e-8, restart=200, maxiter=2000, inner_prec='f64',
          outer_prec='f64', inner_reduction=None, callback=None):
    """Restarted GMRES on the device (see GmresState)."""
    return GmresState(apply, b, tol=tol, restart=restart, maxiter=maxiter, inner_prec=inner_prec,
                      outer_prec=outer_prec, inner_reduction=inner_reduction, x0=x0,
                      callback=callback).run()


MODES = {
    # name: (inner_prec, outer_prec, inner_reduction, coarse residual prec)
    'f64': ('f64', 'f64', None, None),
    # (a coarse FP32 residual would repeat the inner operator: none here)
    'mixed': ('f32', 'f64', 1e-4, None),
    # FP32 treecode farfield in the Arnoldi cycles (operators with a tree
    # attached), exact FP64 direct farfield for the residuals that decide;
    # a cycle stops at 3e-3 since the treecode's own error (~1e-3) bounds
    # what one cycle can gain
    'tree': ('tree', 'f64', 3e-3, 'f32'),
}
# residual updates r <- r - A d in FP32 (direct farfield, FP32 nearfield
# blocks) once the FP64 residual is below 1e-3 ||b||: FP32's ~1e-6 relative
# error on A d stays far below the 1e-8 target
UPDATE_PREC = {'f64': None, 'mixed': 'f32', 'tree': 'f32'}


def make_state(op, b, mode='mixed', tol=1e-8, restart=200, maxiter=2000, callback=None,
               inner_reduction=None, step_callback=None):
    inner, outer, red, coarse = MODES[mode]
    return GmresState(op.apply, b, tol=tol, restart=restart, maxiter=maxiter, inner_prec=inner,
                      outer_prec=outer, inner_reduction=inner_reduction or red, callback=callback,
                      coarse_prec=coarse, update_prec=UPDATE_PREC[mode], operator=op,
                      step_callback=step_callback)


def solve_constrained(A, cm, rhs_dev, tol=1e-8, mode='mixed', restart=1500, maxiter=6000,
                      callback=None, step_callback=None):
    """Solve cm^T A cm x = cm^T rhs; return (u = cm x (device), GmresResult).

    mode 'f64': the plain FP64 path (FP64 farfield everywhere).
    mode 'mixed': FP32 farfield in the Arnoldi cycles, FP64 residual at
    every restart (GMRES-based iterative refinement).
    mode 'tree': as 'mixed' with the FP32 treecode farfield in the cycles
    (A's farfield must have a tree attached); residuals stay FP64 direct."""
    if mode not in MODES:
        raise ValueError('mode must be one of %s' % sorted(MODES))
    op = ConstrainedOperator(A, cm)
    b = op.cmT.dot_dev(rhs_dev)
    res = make_state(op, b, mode, tol, restart, maxiter, callback,
                     step_callback=step_callback).run()
    u = op.cm.dot_dev(res.x)
    res.n_matvec = dict(op.n_matvec)
    return u, res


def iterative_solve(iop, constraints, rhs=None, tol=1e-8, mode='mixed', restart=1500):
    """Host-array API of examples/solve.py:25-61: constraints -> cm, solve
    cm^T (rhs - A c_rhs) and return (cm x + c_rhs as numpy, GmresResult)."""
    from .constraints import build_constraint_matrix
    cm, c_rhs, _ = build_constraint_matrix(constraints, iop.shape[1])
    dev = iop.device
    c_rhs_d = torch.as_tensor(c_rhs, dtype=torch.float64, device=dev)
    Ac = iop.dot_dev(c_rhs_d, None, 'f64')
    if rhs is None:
        r = -Ac
    else:
        r = torch.as_tensor(np.asarray(rhs, dtype=np.float64), device=dev) - Ac
    u, res = solve_constrained(iop, cm, r, tol=tol, mode=mode, restart=restart)
    return (u + c_rhs_d).cpu().numpy(), res
