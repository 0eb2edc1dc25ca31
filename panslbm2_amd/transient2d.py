"""Transient natural-convection heat sink with its adjoint run backward in time
(production/heatsink_transient.cpp:138-222), on the device.

The forward problem runs nt - 1 steps of AD::MacroBrinkmanCollideNaturalConvection from the
uniform state; the objective is the heat-source edge temperature (j = 0, i < L) averaged over all
nt steps.  The adjoint starts from the last velocity and runs AAD::MacroBrinkmanCollideNaturalConvection
backward from step nt - 2 to 0, each step reading that step's forward fields and accumulating
AAD::SensitivityTemperatureAtHeatSource into dfds.  Boundaries are the program's: mirror at x = 0
and bounce-back elsewhere for the flow, T = tem0 on x = lx - 1 and y = ly - 1, the heat flux qn into
y = 0 for x < L, adiabatic elsewhere on y = 0.

Trajectory storage in HBM.  A step's forward fields are rho, ux, uy, tem and the temperature
populations' snapshot gi (13 N doubles on D2Q9).  storage="full" keeps all nt of them resident,
as the reference keeps them in host memory (heatsink_transient.cpp:48-54); the reference's own
141 x 161 grid over 1e5 steps needs 236 GB, which one MI355X's 288 GB holds.  storage="checkpoint"
keeps every K-th step (the two lattices' populations and that step's fields, K = ceil(sqrt(nt))
by default) and recomputes the K steps of each window on a second pair of lattices while the
adjoint walks back through it: one more forward pass, and 31 N nt / K + 13 N (K + 1) doubles.
Both give the same values: the recomputed steps repeat the same launches on the same inputs.

Each step is one fused launch (stream, boundary program, corner smoothing and collide; the
boundary calls read the previous step's velocity through the program's field binding), and the
adjoint step adds the sensitivity launches.  use_graphs=True (trajectory resident, nt even)
captures chunks of steps once as HIP graphs and replays them in every later evaluation (the slots'
addresses do not change; a new design changes only the coefficient arrays' contents); on the
program's 141 x 161 grid it measured the same time as plain launches, the step being device time
(a latency-bound fused launch plus the sensitivity kernels), so it is off by default.  FP64
storage (the reference's transient is FP64).

Differences from the reference program: the design s is taken as given (the program builds it with
its Heaviside filter, available as lbm_heaviside_filter); the snapshot of step 0, which the
program never writes before its adjoint's last step reads it, holds the initial populations.
"""
import ctypes
import math

import numpy as np

from . import D, D2Q9, NC, Context, Lattice, check, hptr, to_face

VP = ctypes.c_void_p


class TransientHeatsink2D:
    def __init__(self, ctx=None, lx=141, ly=161, nt=1000, s=None, seed=0, nu=0.1, pr=6.0, ra=2e5, qn=1e-2, L=4.0,
                 alpha_max=1e4, qf=1e-2, qg=1.0, tem0=0.0, gx=0.0, gy=None, storage="full", window=None,
                 use_graphs=False, chunk=100):
        if storage not in ("full", "checkpoint"):
            raise ValueError("storage is 'full' or 'checkpoint'")
        self.ctx = ctx or Context(0)
        self.Lb = self.ctx.lib
        c = self.ctx
        self.lx, self.ly, self.N, self.nt = lx, ly, lx * ly, nt
        self.nu, self.qn, self.Ledge, self.tem0 = nu, qn, L, tem0
        N = self.N
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        if s is None:
            s = np.random.default_rng(seed).uniform(0.01, 0.99, N)
        s = np.asarray(s, dtype=np.float64).reshape(N)
        # heatsink_transient.cpp:46 (U, diffusivities, gravity) and :108-113 (coefficients)
        u = nu * math.sqrt(ra / pr) / (ly - 1.0)
        self.gx, self.gy = gx, (u * u / (ly - 1.0) if gy is None else gy)
        dfl = nu / pr
        dso = 10.0 * dfl
        self._coef_par = (alpha_max, qf, qg, dfl, dso)
        self.d = {k: c.array((N,), data=v) for k, v in self._coefficients(s).items()}
        for k in ("qx", "qy", "ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy", "dfds"):
            self.d[k] = c.array((N,), fill=0.0)
        self.d["ig"] = c.array((NC[D2Q9], N), fill=0.0)

        F = lambda a, dt=np.float64: to_face(D2Q9, lx, ly, 1, a, dt)
        self.faces = dict(
            type_f=F(np.where(i == 0, 2, 1), np.int32), type_g=F(np.where(i == 0, 2, 0), np.int32),
            set_t=F((i == lx - 1) | (j == ly - 1), np.int32), bottom=F(j == 0, np.int32),
            heat_edge=F((j == 0) & (i < L), np.int32), qn=F(np.where((j == 0) & (i < L), qn, 0.0)),
            tem0=F(np.full(N, tem0)))
        self.edge = np.nonzero((j == 0) & (i < L))[0]
        assert np.array_equal(self.edge, np.arange(len(self.edge))), "edge nodes lead the node order"

        # trajectory: window slots (all nt steps when full), checkpoints when checkpointing
        self.storage = storage
        self.K = nt if storage == "full" else max(2, int(window or math.ceil(math.sqrt(nt))))
        W = nt if storage == "full" else self.K + 1
        self.slots = {k: c.array((W, N), fill=0.0) for k in ("rho", "ux", "uy", "tem")}
        self.slots["gi"] = c.array((W, NC[D2Q9] * N), fill=0.0)
        self.nck = 0 if storage == "full" else (nt - 1) // self.K + 1
        if self.nck:
            self.ck_pops = c.array((self.nck, 2 * NC[D2Q9] * N), fill=0.0)
            self.ck_fields = c.array((self.nck, 13 * N), fill=0.0)
        self.edge_t = c.array((nt, len(self.edge)), fill=0.0)

        # forward lattices (the recompute pair in checkpoint mode) and adjoint lattices
        self.fa, self.fb = Lattice(c, D2Q9, lx, ly), Lattice(c, D2Q9, lx, ly)
        self.pa, self.pb = Lattice(c, D2Q9, lx, ly), Lattice(c, D2Q9, lx, ly)
        self.fwd_prog = self._record_forward()
        self.adj_prog = self._record_adjoint()
        self.objective = None
        # a replayed graph expects each lattice's current buffer to be the one it was captured with:
        # each pass makes nt - 2 buffer swaps (its first step collides in place), so an evaluation
        # leaves the buffers where it found them when nt is even
        self.use_graphs, self.chunk = use_graphs and storage == "full" and nt % 2 == 0, chunk
        self._graphs = {}

    def _coefficients(self, s):
        """heatsink_transient.cpp:108-113: Brinkman and diffusivity coefficients of design s."""
        alpha_max, qf, qg, dfl, dso = self._coef_par
        ly = self.ly
        return dict(
            alpha=alpha_max / (ly - 1.0) * qf * (1.0 - s) / (s + qf),
            dads=-alpha_max / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2,
            diff=dso + (dfl - dso) * s * (1.0 + qg) / (s + qg),
            dkds=(dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2,
        )

    def set_design(self, s):
        """A new design: the four coefficient arrays are rewritten in place, so captured graphs
        (which hold their addresses) stay valid."""
        s = np.asarray(s, dtype=np.float64).reshape(self.N)
        for k, v in self._coefficients(s).items():
            self.d[k].upload(v)

    # ---- addresses --------------------------------------------------------------------------
    def _at(self, arr, row, width=None):
        width = self.N if width is None else width
        return VP(arr.ptr.value + 8 * row * width)

    def _slot(self, name, t):
        r = t if self.storage == "full" else t % (self.K + 1)  # a window's K + 1 steps are distinct slots
        return self._at(self.slots[name], r, NC[D2Q9] * self.N if name == "gi" else self.N)

    def _copy(self, dst, src, doubles):
        check(self.Lb.lbm_device_copy(self.ctx.h, dst, src, 8 * doubles))

    # ---- programs ---------------------------------------------------------------------------
    def _record_forward(self):
        L, fc, d = self.Lb, self.faces, self.d
        ux, uy = self._slot("ux", 0), self._slot("uy", 0)
        check(L.lbm_record_begin(self.ctx.h))
        check(L.lbm_stream(self.fa.h))
        check(L.lbm_stream(self.fb.h))
        check(L.lbm_bc(self.fa.h, hptr(fc["type_f"])))
        check(L.lbm_ad_set_t(self.fb.h, hptr(fc["set_t"]), hptr(fc["tem0"]), ux, uy, None))
        check(L.lbm_ad_set_q(self.fb.h, hptr(fc["bottom"]), hptr(fc["qn"]), ux, uy, None, d["diff"].ptr, D(0.0)))
        check(L.lbm_bc(self.fb.h, hptr(fc["type_g"])))
        check(L.lbm_smooth_corner(self.fa.h))
        check(L.lbm_smooth_corner(self.fb.h))
        prog = VP()
        check(L.lbm_record_end(self.ctx.h, ctypes.byref(prog)))
        return prog

    def _record_adjoint(self):
        L, fc = self.Lb, self.faces
        ux, uy = self._slot("ux", 0), self._slot("uy", 0)
        check(L.lbm_record_begin(self.ctx.h))
        check(L.lbm_istream(self.pa.h))
        check(L.lbm_istream(self.pb.h))
        check(L.lbm_aad_iset_t(self.pb.h, hptr(fc["set_t"]), ux, uy, None))
        check(L.lbm_aad_iset_q(self.pb.h, hptr(fc["bottom"]), ux, uy, None, D(0.0)))
        check(L.lbm_aad_iset_q(self.pb.h, hptr(fc["heat_edge"]), ux, uy, None, D(1.0)))
        check(L.lbm_ibc(self.pb.h, hptr(fc["type_g"])))
        check(L.lbm_ibc(self.pa.h, hptr(fc["type_f"])))
        check(L.lbm_smooth_corner(self.pa.h))
        check(L.lbm_smooth_corner(self.pb.h))
        prog = VP()
        check(L.lbm_record_end(self.ctx.h, ctypes.byref(prog)))
        return prog

    def _bind(self, prog, t):
        ux, uy = self._slot("ux", t), self._slot("uy", t)
        for k in range(self.Lb.lbm_program_num_field_sets(prog)):
            check(self.Lb.lbm_program_bind_fields(prog, k, None, ux, uy, None, None, None))

    # ---- steps ------------------------------------------------------------------------------
    def _fwd_step(self, t):
        """Forward step t >= 1: the boundary calls of step t - 1 (reading its velocity), then the
        collide of step t saving into t's slot."""
        L, d = self.Lb, self.d
        prog = None
        if t > 1:
            prog = self.fwd_prog
            self._bind(prog, t - 1)
        check(L.lbm_ad_macro_brinkman_collide_natural_convection_p(
            prog, self.fa.h, self._slot("rho", t), self._slot("ux", t), self._slot("uy", t), None, d["alpha"].ptr,
            D(self.nu), self.fb.h, self._slot("tem", t), d["qx"].ptr, d["qy"].ptr, None, d["diff"].ptr, D(self.gx),
            D(self.gy), D(0.0), D(self.tem0), 1, self._slot("gi", t)))

    def _adj_step(self, t):
        L, d = self.Lb, self.d
        prog = None
        if t < self.nt - 2:
            prog = self.adj_prog
            self._bind(prog, t + 1)
        ux, uy, tem = self._slot("ux", t), self._slot("uy", t), self._slot("tem", t)
        check(L.lbm_aad_macro_brinkman_collide_natural_convection_p(
            prog, self.pa.h, self._slot("rho", t), ux, uy, None, d["ip"].ptr, d["iux"].ptr, d["iuy"].ptr, None,
            d["imx"].ptr, d["imy"].ptr, None, d["alpha"].ptr, D(self.nu), self.pb.h, tem, d["item"].ptr,
            d["iqx"].ptr, d["iqy"].ptr, None, d["diff"].ptr, D(self.gx), D(self.gy), D(0.0), 1, d["ig"].ptr))
        self._sens(t)

    def _sens(self, t):
        L, d = self.Lb, self.d
        check(L.lbm_sensitivity_temperature_at_heat_source(
            self.pb.h, d["dfds"].ptr, self._slot("ux", t), self._slot("uy", t), None, d["imx"].ptr, d["imy"].ptr, None,
            d["dads"].ptr, self._slot("tem", t), d["item"].ptr, d["iqx"].ptr, d["iqy"].ptr, None, self._slot("gi", t),
            d["ig"].ptr, d["diff"].ptr, d["dkds"].ptr, hptr(self.faces["heat_edge"]), hptr(self.faces["qn"])))

    def _run(self, key, steps):
        """Run the callables in steps; with graphs, as one graph captured the first time."""
        if not self.use_graphs:
            for s in steps:
                s()
            return
        g = self._graphs.get(key)
        if g is None:
            check(self.Lb.lbm_graph_begin(self.ctx.h))
            try:
                for s in steps:
                    s()
            except BaseException:
                # end the capture so the context stream works again, and drop the partial graph
                part = VP()
                if self.Lb.lbm_graph_end(self.ctx.h, ctypes.byref(part)) == 0 and part:
                    self.Lb.lbm_graph_destroy(part)
                raise
            g = VP()
            check(self.Lb.lbm_graph_end(self.ctx.h, ctypes.byref(g)))
            self._graphs[key] = g
        check(self.Lb.lbm_graph_launch(g))

    # ---- checkpoints ------------------------------------------------------------------------
    def _fields(self):
        return [("rho", 1), ("ux", 1), ("uy", 1), ("tem", 1), ("gi", NC[D2Q9])]

    def _save(self, c, t):
        N, nc = self.N, NC[D2Q9]
        pops = self._at(self.ck_pops, c, 2 * nc * N)
        self._copy(pops, VP(self.Lb.lbm_lattice_populations(self.fa.h)), nc * N)
        self._copy(VP(pops.value + 8 * nc * N), VP(self.Lb.lbm_lattice_populations(self.fb.h)), nc * N)
        off = self._at(self.ck_fields, c, 13 * N).value
        for name, w in self._fields():
            self._copy(VP(off), self._slot(name, t), w * N)
            off += 8 * w * N

    def _restore(self, c, t):
        N, nc = self.N, NC[D2Q9]
        pops = self._at(self.ck_pops, c, 2 * nc * N)
        self._copy(VP(self.Lb.lbm_lattice_populations(self.fa.h)), pops, nc * N)
        self._copy(VP(self.Lb.lbm_lattice_populations(self.fb.h)), VP(pops.value + 8 * nc * N), nc * N)
        off = self._at(self.ck_fields, c, 13 * N).value
        for name, w in self._fields():
            self._copy(self._slot(name, t), VP(off), w * N)
            off += 8 * w * N

    # ---- the evaluation ---------------------------------------------------------------------
    def _init_forward(self):
        L, N = self.Lb, self.N
        check(L.lbm_device_fill(self.ctx.h, self._slot("rho", 0), D(1.0), N))
        for k in ("ux", "uy", "tem"):
            check(L.lbm_device_fill(self.ctx.h, self._slot(k, 0), D(0.0), N))
        check(L.lbm_device_fill(self.ctx.h, self.d["qx"].ptr, D(0.0), N))
        check(L.lbm_device_fill(self.ctx.h, self.d["qy"].ptr, D(0.0), N))
        check(L.lbm_ns_initial_condition(self.fa.h, self._slot("rho", 0), self._slot("ux", 0), self._slot("uy", 0),
                                         None))
        check(L.lbm_ad_initial_condition(self.fb.h, self._slot("tem", 0), self._slot("ux", 0), self._slot("uy", 0),
                                         None))
        self._copy(self._slot("gi", 0), VP(L.lbm_lattice_populations(self.fb.h)), NC[D2Q9] * N)

    def forward(self):
        """The forward pass; fills the trajectory (full) or the checkpoints, and the objective."""
        ne = len(self.edge)
        self._init_forward()
        if self.nck:
            self._save(0, 0)
        self._copy(self._at(self.edge_t, 0, ne), self._slot("tem", 0), ne)

        def step(t):
            def run():
                self._fwd_step(t)
                self._copy(self._at(self.edge_t, t, ne), self._slot("tem", t), ne)
                if self.nck and t % self.K == 0:
                    self._save(t // self.K, t)
            return run

        for t0 in range(1, self.nt, self.chunk):
            self._run(("f", t0), [step(t) for t in range(t0, min(t0 + self.chunk, self.nt))])
        edge = self.edge_t.download()
        f = 0.0
        for t in range(self.nt):  # heatsink_transient.cpp:196-207, in its order
            for v in edge[t]:
                f += v
        self.objective = f / (self.Ledge * self.nt)
        return self.objective

    def adjoint(self):
        """The adjoint pass backward in time, accumulating dfds (the unfiltered sensitivity)."""
        L, d, nt = self.Lb, self.d, self.nt
        if self.use_graphs and nt >= 2:
            # the heat-source edge operators are built on the first call: make it here, outside any
            # capture (dfds is cleared just below)
            self._sens(0)
        for k in ("ip", "iux", "iuy", "item", "iqx", "iqy", "dfds"):
            d[k].fill(0.0)
        ux, uy = self._slot("ux", nt - 1), self._slot("uy", nt - 1)
        check(L.lbm_ans_initial_condition(self.pa.h, ux, uy, None, d["ip"].ptr, d["iux"].ptr, d["iuy"].ptr, None))
        check(L.lbm_aad_initial_condition(self.pb.h, ux, uy, None, d["item"].ptr, d["iqx"].ptr, d["iqy"].ptr, None))
        if not self.nck:
            for t0 in range(nt - 2, -1, -self.chunk):
                self._run(("a", t0), [(lambda t: (lambda: self._adj_step(t)))(t)
                                      for t in range(t0, max(t0 - self.chunk, -1), -1)])
            return d["dfds"]
        for c in range(self.nck - 1, -1, -1):
            t0 = c * self.K
            t1 = min(t0 + self.K, nt - 1)
            if t1 <= t0:
                continue
            self._restore(c, t0)
            for t in range(t0 + 1, t1 + 1):
                self._fwd_step(t)
            for t in range(t1 - 1, t0 - 1, -1):
                self._adj_step(t)
        return d["dfds"]

    def evaluate(self):
        self.forward()
        self.adjoint()
        self.ctx.sync()
        return self.objective, self.d["dfds"].download()

    def close(self):
        for g in self._graphs.values():
            self.Lb.lbm_graph_destroy(g)
        self._graphs = {}
        for p in (self.fwd_prog, self.adj_prog):
            if p:
                self.Lb.lbm_program_destroy(p)
        self.fwd_prog = self.adj_prog = None
        for lat in (self.fa, self.fb, self.pa, self.pb):
            lat.close()
