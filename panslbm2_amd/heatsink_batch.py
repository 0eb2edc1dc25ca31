"""Many independent 2D heat-sink design evaluations solved together: every time step of every
instance runs in one kernel launch (lbm_batch), 100-step chunks are replayed as HIP graphs, and
the residuals of all instances are reduced in one launch pair (lbm_residual_batch).

Each instance is a panslbm2_amd.heatsink2d.Heatsink2D (its own grid, design, heat source, inlet
velocity, fields and recorded boundary programs); this driver replaces its per-instance time loop.
The step order per instance is exactly the single-instance driver's, so with FP64 storage every
instance's result is value-identical to solving it alone.

Precision as in Heatsink2D: with prec=F32 all instances run in FP32 until every active one's
residual has reached the FP32 floor or has stopped falling (not 5 % lower in ten checks below
100x the floor, or in twenty below STALL_CEILING), then every lattice converts to FP64 and the
solves finish at eps.
An instance that has converged is frozen (its workgroups return at once) at that check step, so
each instance stops where it would have stopped alone.  Instances whose reduced-precision solve
diverges are solved again from their cold start in FP64, together in one more batched phase.
"""
import ctypes
import time

import numpy as np

from . import D2Q5, F32, F64, Context, check
from .heatsink2d import STALL_CEILING, Heatsink2D

VP = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


class HeatsinkBatch2D:
    def __init__(self, ctx=None, specs=(), ad_kind=D2Q5, prec=F32, eps=1e-6, check_every=100, max_steps=400000,
                 use_graphs=True, tune=True, fp32_floor=None, filter_radius=3.0):
        """specs: one dict per instance with lx, ly and optionally seed, heat, umax."""
        self.ctx = ctx or Context(0)
        self.L = self.ctx.lib
        self.prec, self.eps, self.ce, self.max_steps = prec, eps, check_every, max_steps
        self.use_graphs, self.tune = use_graphs, tune
        self.hs = [Heatsink2D(self.ctx, lx=sp["lx"], ly=sp["ly"], ad_kind=ad_kind, seed=sp.get("seed", i),
                              heat=sp.get("heat", 1e-3), umax=sp.get("umax", 0.05), eps=eps, check_every=check_every,
                              max_steps=max_steps, use_graphs=False, prec=prec, filter_radius=filter_radius,
                              fp32_floor=fp32_floor)
                   for i, sp in enumerate(specs)]
        self.fp32_floor = self.hs[0].fp32_floor if self.hs else fp32_floor
        self.n = len(self.hs)
        self.N = np.array([h.N for h in self.hs])
        self.tuned = None

    # ---- batches of one step kind ----------------------------------------------------------------
    def _collect(self, step_fn, t):
        check(self.L.lbm_batch_begin(self.ctx.h))
        try:
            for h in self.hs:
                step_fn(h, t)
        finally:
            b = VP()
            rc = self.L.lbm_batch_end(self.ctx.h, ctypes.byref(b))
        check(rc)
        return b

    def _make_batches(self, step_fn, t_next):
        """Batches for the four kinds of step (parity x save mode), collected in alternating parity
        starting with that of step t_next so that the lattices' buffer bookkeeping returns to where
        it was.  Key: (t % 2, save mode)."""
        ce = self.ce
        base = (t_next // ce) * ce
        reps = {}  # a representative step of each kind, > 1, in [base + 1, base + 2 ce]
        for t in range(base + 1, base + 2 * ce + 1):
            if t < 2:
                continue
            key = (t % 2, self.hs[0]._save_mode(t, self.adjoint))
            reps.setdefault(key, t)
        order = sorted(reps.items(), key=lambda kv: (kv[0][0] != t_next % 2, kv[0][1]))
        even = [kv for kv in order if kv[0][0] == 0]
        odd = [kv for kv in order if kv[0][0] == 1]
        first, second = (even, odd) if t_next % 2 == 0 else (odd, even)
        seq = [x for pair in zip(first, second) for x in pair]
        return {key: self._collect(step_fn, t) for key, t in seq}

    def _set_active(self, batches, active):
        flags = np.ascontiguousarray(active, dtype=np.int32)
        for b in batches.values():
            check(self.L.lbm_batch_set_active(b, flags.ctypes.data_as(VP)))

    def _destroy(self, batches):
        for b in batches.values():
            self.L.lbm_batch_destroy(b)

    def _launch(self, batches, t):
        check(self.L.lbm_batch_launch(batches[(t % 2, self.hs[0]._save_mode(t, self.adjoint))]))

    def _residuals(self, t):
        """Per instance: max of the two residuals after step t."""
        n = self.n
        out = np.zeros(n)
        worst = np.zeros(n)
        for k in range(2):
            a = (VP * (3 * n))()
            b = (VP * (3 * n))()
            for i, h in enumerate(self.hs):
                f = h.residual_fields(self.adjoint, t)[k]
                a[3 * i], a[3 * i + 1] = f[0].ptr, f[1].ptr
                b[3 * i], b[3 * i + 1] = f[2].ptr, f[3].ptr
            lens = (I * n)(*[int(x) for x in self.N])
            check(self.L.lbm_residual_batch(self.ctx.h, n, 2, a, b, lens, out.ctypes.data_as(ctypes.POINTER(D))))
            worst = np.maximum(worst, out)
        return worst

    def _set_prec(self, prec):
        for h in self.hs:
            h._set_prec(prec)

    # ---- one phase (forward or adjoint) ----------------------------------------------------------
    def _configure(self, batches):
        if self.tuned is not None:
            for b in batches.values():
                check(self.L.lbm_batch_configure_tiles(b, *self.tuned))

    def _phase(self, step_fn):
        """Steps until every instance's residuals are below eps; returns (steps, convergence step per
        instance: -1 not converged by max_steps, -2 diverged).  Residuals are checked after steps
        ce, 2 ce, ... as in the single-instance driver."""
        ce = self.ce
        for h in self.hs:  # step 1 has no recorded program: run it per instance
            step_fn(h, 1)
        t = 1
        batches = self._make_batches(step_fn, 2)
        if self.tune and self.tuned is None:
            quiet = self.hs[0]._save_mode(2, self.adjoint)  # save mode of the steps between checks
            arr = (VP * 2)(batches[(0, quiet)], batches[(1, quiet)])
            blk, rows, xcd, rounds = I(), I(), I(), I()
            check(self.L.lbm_batch_tune_tiles(arr, 2, 3, ctypes.byref(blk), ctypes.byref(rows), ctypes.byref(xcd),
                                              ctypes.byref(rounds)))
            self.tuned = (blk.value, rows.value, xcd.value)  # workgroup size, tile rows, XCD remap
            # the tuner's rounds are real steps, of the kind between checks; the residual checks
            # resume at the next multiple of check_every after them
            t += 2 * rounds.value
        else:
            self._configure(batches)
        while t % ce:  # finish the first chunk
            t += 1
            self._launch(batches, t)
        graph = None
        conv = np.full(self.n, -1)
        best, stall = np.full(self.n, np.inf), np.zeros(self.n, dtype=int)
        while True:
            r = self._residuals(t)
            newly = (conv == -1) & (r < self.eps)
            conv[newly] = t
            bad = (conv == -1) & ~np.isfinite(r)  # a diverged instance stops and is reported (-2)
            conv[bad] = -2
            newly |= bad
            if np.all(conv != -1) or t >= self.max_steps:
                break
            if np.any(newly):  # freeze at this check step (even: buffer bookkeeping matches)
                self._set_active(batches, conv == -1)
            act = conv == -1
            if self.hs[0].pf.prec == F32:
                stall = np.where(r < 0.95 * best, 0, stall + 1)
                best = np.minimum(best, r)
                ready = ((r < self.fp32_floor) | ((stall >= 10) & (r < 100 * self.fp32_floor))
                         | ((stall >= 20) & (r < STALL_CEILING)))
                if np.all(ready[act]):
                    self.t32 = t
                    if graph is not None:
                        self.L.lbm_graph_destroy(graph)
                        graph = None
                    self._destroy(batches)
                    self._set_prec(F64)
                    batches = self._make_batches(step_fn, t + 1)
                    self._configure(batches)
                    self._set_active(batches, conv == -1)
            # steps t+1 .. t+ce: the same launch sequence for every chunk (ce is even)
            if not self.use_graphs:
                for s in range(t + 1, t + ce + 1):
                    self._launch(batches, s)
            else:
                if graph is None:
                    check(self.L.lbm_graph_begin(self.ctx.h))
                    for s in range(t + 1, t + ce + 1):
                        self._launch(batches, s)
                    graph = VP()
                    check(self.L.lbm_graph_end(self.ctx.h, ctypes.byref(graph)))
                check(self.L.lbm_graph_launch(graph))
            t += ce
        if graph is not None:
            self.L.lbm_graph_destroy(graph)
        self._destroy(batches)
        return t, conv

    def _record(self, conv, t_end, kind, hs=None):
        """Each instance's step count and buffer parity: a frozen instance's state is that of its
        convergence step, the others' that of the phase's last step."""
        for h, c in zip(hs if hs is not None else self.hs, conv):
            steps = int(c) if c > 0 else t_end
            setattr(h, kind + "_steps", steps)
            setattr(h, kind + "_buf", steps % 2)

    def _redo_in_fp64(self, conv, step_fn, init, kind):
        """Instances whose reduced-precision solve diverged (conv -2) are solved again from their
        cold start in FP64, together in one batch; conv is updated in place.  Returns their
        indices."""
        idx = [i for i, c in enumerate(conv) if c == -2] if self.prec != F64 else []
        if not idx:
            return idx
        saved = (self.hs, self.n, self.N, self.prec)
        sub = [self.hs[i] for i in idx]
        precs = [h.prec for h in sub]
        try:
            for h in sub:
                h.prec = F64
                init(h)
            self.hs, self.n, self.N, self.prec = sub, len(sub), np.array([h.N for h in sub]), F64
            t, c2 = self._phase(step_fn)
        finally:
            for h, p in zip(sub, precs):
                h.prec = p
            self.hs, self.n, self.N, self.prec = saved
        self._record(c2, t, kind, sub)
        for i, c in zip(idx, c2):
            conv[i] = c
        return idx

    def evaluate(self, warm=False):
        """Forward, adjoint and sensitivity of every instance; returns timings and step counts.
        warm=True starts each instance's solves from its last converged solution (Heatsink2D
        keep_solution, done after every converged evaluation) where it has one; an instance whose
        solve diverges is solved again from its cold start in FP64."""
        t0 = time.perf_counter()
        self.adjoint = False
        self.t32 = 0
        for h in self.hs:
            h.init_forward(warm)
        fwd_step = lambda h, t: h._fwd_step(t)
        tf, conv_f = self._phase(fwd_step)
        t32f = self.t32
        self._record(conv_f, tf, "fwd")
        redo_f = self._redo_in_fp64(conv_f, fwd_step, lambda h: h.init_forward(), "fwd")
        self.adjoint = True
        self.t32 = 0
        for h in self.hs:
            h.init_adjoint(warm)
        adj_step = lambda h, t: h._adj_step(t, h.fwd_buf)
        ta, conv_a = self._phase(adj_step)
        t32a = self.t32
        self._record(conv_a, ta, "adj")
        redo_a = self._redo_in_fp64(conv_a, adj_step, lambda h: h.init_adjoint(), "adj")
        self.t32 = t32a
        objectives = [h.sensitivity() for h in self.hs]
        for h, cf, ca in zip(self.hs, conv_f, conv_a):
            if cf > 0 and ca > 0:
                h.keep_solution()
        self.ctx.sync()
        dt = time.perf_counter() - t0
        return {"seconds": dt, "fwd_steps": tf, "adj_steps": ta, "fwd_conv": conv_f.tolist(),
                "adj_conv": conv_a.tolist(), "fp32_steps": (t32f, self.t32), "objectives": objectives,
                "fp64_redo": (redo_f, redo_a),
                "tuned": self.tuned,
                "node_steps": int(sum(h.N * (h.fwd_steps + h.adj_steps) for h in self.hs))}

    def close(self):
        for h in self.hs:
            h.close()
