"""2D forced-convection heatsink design evaluation on the HIP path: forward steady solve, adjoint
steady solve, design sensitivity.

Step order and convergence test follow production/heatsink.cpp:136-232 (collide with saving,
Residual every `check` steps on the step-to-step change of u and q, stream, boundary calls,
SmoothCorner, swap of the saved-velocity arrays), with the forced-convection collides of
advection.h:814 and adjointadvection.h:997 and the boundary set of test/nsadsens.cpp:79-131.

Workload (one design evaluation):
  * NS on D2Q9 with Brinkman drag alpha = alpha_max/(ly-1)*qf*(1-gamma)/(qf+gamma), the
    reference's scaling (heatsink.cpp:110; with alpha_max = 1e4 unscaled the reference's adjoint
    Brinkman terms diverge within a few hundred steps); AD on D2Q5 (or on
    D2Q9, the reference's own pairing) with diffusivity D(gamma) interpolated between fluid
    D_f = nu/Pr and solid 10*D_f (heatsink.cpp:46,109) and a volumetric heat source
    Q*(1-gamma) in the solid.
  * Parabolic inlet (u_max) on x = 0, pressure outlet (rho = 1) on x = lx-1, bounce-back walls
    on y = 0 and y = ly-1; T = 0 at the inlet, zero heat flux on walls and outlet.
  * gamma: seeded uniform[0,1] smoothed by the density filter (densityfilter.h:512-515).
  * Objective: mean temperature on the bottom-wall segment x < lx/2 (heatsink.cpp:215-226),
    injected into the adjoint through the flux condition with eps = 1 (heatsink.cpp:202).

Every time step is one fused kernel pass (pull stream + boundary program + corner smoothing +
collide); 100-step chunks between residual checks run as one HIP graph.
"""
import ctypes
import time

import numpy as np

from . import D2Q5, D2Q9, F16, F32, F64, NC, Context, Lattice, check, hptr, to_face

D = ctypes.c_double
VP = ctypes.c_void_p


# A reduced-precision solve whose residual has stopped falling (not 5 % lower in ten checks) below
# 100x the switch floor is at that precision's floor; between that and STALL_CEILING (the FP32
# heat-flux residual of a high-Peclet channel levels off at ~1e-5) it must stall twice as long, so
# that a slow stretch of the convergence is not taken for the floor; plateaus of the early
# transient lie above STALL_CEILING.
STALL_CEILING = 1e-4


def switch_ready(worst, floor, stall, ceiling=STALL_CEILING):
    """Whether a reduced-precision solve converts to FP64: its residual is below the switch floor,
    or has stalled (no 5 % drop) for ten checks below 100 floor or twenty below ceiling."""
    return worst < floor or (stall >= 10 and worst < 100.0 * floor) or (stall >= 20 and worst < ceiling)


class Heatsink2D:
    stall_ceiling = STALL_CEILING  # see switch_ready
    def __init__(self, ctx=None, lx=200, ly=400, ad_kind=D2Q5, seed=0, nu=0.1, pr=6.0, umax=0.05, heat=1e-3,
                 alpha_max=1e4, qf=0.01, qg=1.0, filter_radius=3.0, eps=1e-6, check_every=100, max_steps=400000,
                 use_graphs=True, prec=F64, fp32_floor=None):
        self.ctx = ctx or Context(0)
        self.L = self.ctx.lib
        self.lx, self.ly, self.N = lx, ly, lx * ly
        self.ad_kind = ad_kind
        self.nu, self.eps, self.check_every, self.max_steps = nu, eps, check_every, max_steps
        assert check_every % 2 == 0, "chunks must leave the alternating buffers where they found them"
        self.use_graphs = use_graphs
        N = self.N
        c = self.ctx
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        self.i, self.j = i, j

        # design field: seeded uniform[0,1], density filter on the device
        rng = np.random.default_rng(seed)
        self.x = rng.uniform(0.0, 1.0, N)
        raw = c.array((N,), data=self.x)
        gam = c.array((N,), fill=0.0)
        check(self.L.lbm_density_filter(c.h, lx, ly, 1, D(filter_radius), raw.ptr, gam.ptr))
        s = gam.download()
        self.gamma = s
        self.alpha_max, self.qf, self.qg, self.heat, self.filter_radius = alpha_max, qf, qg, heat, filter_radius
        self.dfl = nu / pr
        self.dso = 10.0 * self.dfl
        self.d = {k: c.array((N,), data=v) for k, v in self._coefficients(s).items()}
        self.d["raw"], self.d["design"] = raw, gam
        self.kept = None  # device copies of the last converged fields (the warm start)
        # fields: saved values alternate between two buffers per quantity (u at t and t-1)
        for k in ("rho", "tem", "ip", "imx", "imy", "item", "dfds"):
            self.d[k] = c.array((N,), fill=0.0)
        for k in ("ux", "uy", "qx", "qy", "iux", "iuy", "iqx", "iqy"):
            self.d[k + "0"] = c.array((N,), fill=0.0)
            self.d[k + "1"] = c.array((N,), fill=0.0)
        self.d["g"] = c.array((NC[ad_kind], N), fill=0.0)
        self.d["ig"] = c.array((NC[ad_kind], N), fill=0.0)
        self.d["one"] = c.array((N,), fill=1.0)
        self.d["zero"] = c.array((N,), fill=0.0)

        # prec: population storage of both lattices.  F32 / F16: each solve runs in the reduced
        # precision until its residual reaches fp32_floor or stops falling, then converts the
        # lattices to FP64 and finishes at eps (mixed precision).  Default floors: 1e-7 (F32,
        # below the usual eps: FP32 runs until it converges or its residual stalls), 1e-3 (F16,
        # whose step-to-step residual levels off at a few 1e-3).
        self.prec = prec
        self.fp32_floor = fp32_floor if fp32_floor is not None else (1e-3 if prec == F16 else 1e-7)
        self.pf = Lattice(c, D2Q9, lx, ly, prec=prec)
        self.pg = Lattice(c, ad_kind, lx, ly, prec=prec)

        # boundary lambdas, tabulated
        H = ly - 1.0
        uin = 4.0 * umax * j * (H - j) / H ** 2
        inlet = ((i == 0) & (j > 0) & (j < ly - 1)).astype(np.int32)
        outlet = ((i == lx - 1) & (j > 0) & (j < ly - 1)).astype(np.int32)
        walls = ((j == 0) | (j == ly - 1)).astype(np.int32)
        flux = (walls | outlet).astype(np.int32)
        self.obj_nodes = np.nonzero((j == 0) & (i < lx // 2))[0].astype(np.int32)
        objedge = ((j == 0) & (i < lx // 2)).astype(np.int32)
        F = lambda kind, a, dt=np.float64: to_face(kind, lx, ly, 1, a, dt)
        self.faces = dict(
            walls_f=F(D2Q9, walls, np.int32), inlet_f=F(D2Q9, inlet, np.int32), outlet_f=F(D2Q9, outlet, np.int32),
            uin_f=F(D2Q9, uin), zero_f=F(D2Q9, np.zeros(N)), one_f=F(D2Q9, np.ones(N)),
            inlet_g=F(ad_kind, inlet, np.int32), flux_g=F(ad_kind, flux, np.int32), obj_g=F(ad_kind, objedge, np.int32),
            zero_g=F(ad_kind, np.zeros(N)), outlet_type=F(D2Q9, outlet * 2, np.int32),
        )
        self.fwd_prog = self._record_forward()
        self.adj_prog = None
        self._graphs = {}

    def _coefficients(self, s):
        """Brinkman, diffusivity and heat-source coefficients of design gamma = s (heatsink.cpp:104-117
        interpolation, heat source Q(1 - gamma)) and their design derivatives."""
        am, qf, qg, dfl, dso, ly = self.alpha_max, self.qf, self.qg, self.dfl, self.dso, self.ly
        return {
            "alpha": am / (ly - 1.0) * qf * (1.0 - s) / (s + qf),
            "dads": -am / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2,
            "diff": dso + (dfl - dso) * s * (1.0 + qg) / (s + qg),
            "dkds": (dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2,
            "heat": self.heat * (1.0 - s),
            "dhds": np.full(self.N, -self.heat),
        }

    def set_design(self, x):
        """Design variables x (per node, clipped to [0, 1]) -> gamma = density-filtered x -> the
        coefficient arrays, rewritten in place (recorded programs and graphs stay valid)."""
        c, d = self.ctx, self.d
        self.x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 1.0).copy()
        d["raw"].upload(self.x)
        check(self.L.lbm_density_filter(c.h, self.lx, self.ly, 1, D(self.filter_radius), d["raw"].ptr,
                                        d["design"].ptr))
        self.gamma = d["design"].download()
        for k, v in self._coefficients(self.gamma).items():
            d[k].upload(v)

    def design_update(self, vf=0.4, move=0.05):
        """One projected-gradient design step over every node, as Heatsink3D's optimizer="pg":
        x <- clip(x - move dF/dgamma / max|dF/dgamma| + lam, x - move, x + move, 0, 1), lam found by
        bisection so that the solid fraction mean(1 - x) equals vf.  dF/dgamma serves as the
        gradient in x (the density filter's chain is not applied)."""
        g = self.d["dfds"].download()
        g = g / max(np.abs(g).max(), 1e-300)
        x = self.x
        lo_b, hi_b = np.maximum(0.0, x - move), np.minimum(1.0, x + move)
        trial = lambda lam: np.clip(x - move * g + lam, lo_b, hi_b)
        lo, hi = -1.0, 1.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if trial(mid).mean() > 1.0 - vf:
                hi = mid
            else:
                lo = mid
        self.set_design(trial(0.5 * (lo + hi)))

    def keep_solution(self):
        """Device copies of the converged forward and adjoint fields: the next warm start."""
        c, d, N = self.ctx, self.d, self.N
        fb, ab = self.fwd_buf, self.adj_buf
        src = {"rho": d["rho"], "ux": d["ux%d" % fb], "uy": d["uy%d" % fb], "tem": d["tem"], "ip": d["ip"],
               "iux": d["iux%d" % ab], "iuy": d["iuy%d" % ab], "item": d["item"], "iqx": d["iqx%d" % ab],
               "iqy": d["iqy%d" % ab]}
        if self.kept is None:
            self.kept = {k: c.array((N,)) for k in src}
        for k, a in src.items():
            check(self.L.lbm_device_copy(c.h, self.kept[k].ptr, a.ptr, 8 * N))

    def _copy(self, dst, src):
        check(self.L.lbm_device_copy(self.ctx.h, dst.ptr, src.ptr, 8 * self.N))

    # ---- programs ---------------------------------------------------------------------------
    def _record_forward(self):
        L, c, d, fc = self.L, self.ctx, self.d, self.faces
        check(L.lbm_record_begin(c.h))
        check(L.lbm_stream(self.pf.h))
        check(L.lbm_stream(self.pg.h))
        check(L.lbm_bc(self.pf.h, hptr(fc["walls_f"])))
        check(L.lbm_ns_set_u(self.pf.h, hptr(fc["inlet_f"]), hptr(fc["uin_f"]), hptr(fc["zero_f"]), hptr(fc["zero_f"])))
        check(L.lbm_ns_set_rho(self.pf.h, hptr(fc["outlet_f"]), hptr(fc["one_f"]), hptr(fc["zero_f"]),
                               hptr(fc["zero_f"])))
        check(L.lbm_smooth_corner(self.pf.h))
        check(L.lbm_ad_set_t(self.pg.h, hptr(fc["inlet_g"]), hptr(fc["zero_g"]), d["ux0"].ptr, d["uy0"].ptr, None))
        check(L.lbm_ad_set_q(self.pg.h, hptr(fc["flux_g"]), hptr(fc["zero_g"]), d["ux0"].ptr, d["uy0"].ptr, None,
                             d["diff"].ptr, D(0.0)))
        check(L.lbm_smooth_corner(self.pg.h))
        prog = VP()
        check(L.lbm_record_end(c.h, ctypes.byref(prog)))
        assert L.lbm_program_fusable(prog) == 1
        return prog

    def _record_adjoint(self, b):
        """b: buffer holding the converged forward velocity."""
        L, c, d, fc = self.L, self.ctx, self.d, self.faces
        ux, uy = d["ux%d" % b].ptr, d["uy%d" % b].ptr
        check(L.lbm_record_begin(c.h))
        check(L.lbm_istream(self.pf.h))
        check(L.lbm_istream(self.pg.h))
        check(L.lbm_aad_iset_t(self.pg.h, hptr(fc["inlet_g"]), ux, uy, None))
        check(L.lbm_aad_iset_q(self.pg.h, hptr(fc["flux_g"]), ux, uy, None, D(0.0)))
        check(L.lbm_aad_iset_q(self.pg.h, hptr(fc["obj_g"]), ux, uy, None, D(1.0)))
        check(L.lbm_smooth_corner(self.pg.h))
        check(L.lbm_ibc(self.pf.h, hptr(fc["walls_f"])))
        check(L.lbm_ans_iset_u(self.pf.h, hptr(fc["inlet_f"]), hptr(fc["uin_f"]), hptr(fc["zero_f"]),
                               hptr(fc["zero_f"]), D(0.0), 1))
        check(L.lbm_aad_iset_rho(self.pf.h, self.pg.h, hptr(fc["outlet_type"]), d["rho"].ptr, ux, uy, None,
                                 d["tem"].ptr, D(0.0)))
        check(L.lbm_smooth_corner(self.pf.h))
        prog = VP()
        check(L.lbm_record_end(c.h, ctypes.byref(prog)))
        assert L.lbm_program_fusable(prog) == 1
        return prog

    # ---- steps ------------------------------------------------------------------------------
    def _save_mode(self, t, adjoint):
        """1 (every field) on a residual step and the step before it; otherwise the forward step
        writes only the boundary nodes' velocity its boundary operators read next (3) and the
        adjoint step nothing."""
        r = t % self.check_every
        if r == 0 or r == self.check_every - 1:
            return 1
        return 0 if adjoint else 3

    def _fwd_step(self, t):
        """Forward step t (t >= 1): saves into buffer t % 2; boundary calls read buffer (t-1) % 2."""
        L, d = self.L, self.d
        b, pb = t % 2, (t - 1) % 2
        prog = None
        if t > 1:
            prog = self.fwd_prog
            for s in range(L.lbm_program_num_field_sets(prog)):
                check(L.lbm_program_bind_fields(prog, s, None, d["ux%d" % pb].ptr, d["uy%d" % pb].ptr, None, None,
                                                None))
        mode = self._save_mode(t, False)
        check(L.lbm_ad_macro_brinkman_collide_force_convection_p(
            prog, self.pf.h, d["rho"].ptr, d["ux%d" % b].ptr, d["uy%d" % b].ptr, None, d["alpha"].ptr, D(self.nu),
            self.pg.h, d["tem"].ptr, d["qx%d" % b].ptr, d["qy%d" % b].ptr, None, d["diff"].ptr, mode, d["g"].ptr,
            d["heat"].ptr))

    def _adj_step(self, t, fb):
        L, d = self.L, self.d
        b = t % 2
        prog = self.adj_prog if t > 1 else None
        mode = self._save_mode(t, True)
        check(L.lbm_aad_macro_brinkman_collide_force_convection_p(
            prog, self.pf.h, d["rho"].ptr, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, None, d["ip"].ptr,
            d["iux%d" % b].ptr, d["iuy%d" % b].ptr, None, d["imx"].ptr, d["imy"].ptr, None, d["alpha"].ptr, D(self.nu),
            self.pg.h, d["tem"].ptr, d["item"].ptr, d["iqx%d" % b].ptr, d["iqy%d" % b].ptr, None, d["diff"].ptr, mode,
            d["ig"].ptr))

    def _chunk(self, key, t0, step):
        """Steps t0+1 .. t0+check_every, captured once as a HIP graph (t0 even, so bindings repeat)."""
        if not self.use_graphs:
            for t in range(t0 + 1, t0 + self.check_every + 1):
                step(t)
            return
        g = self._graphs.get(key)
        if g is None:
            check(self.L.lbm_graph_begin(self.ctx.h))
            try:
                for t in range(t0 + 1, t0 + self.check_every + 1):
                    step(t)
            except BaseException:
                # end the capture so the context stream works again, and drop the partial graph
                part = VP()
                if self.L.lbm_graph_end(self.ctx.h, ctypes.byref(part)) == 0 and part:
                    self.L.lbm_graph_destroy(part)
                raise
            g = VP()
            check(self.L.lbm_graph_end(self.ctx.h, ctypes.byref(g)))
            self._graphs[key] = g
        check(self.L.lbm_graph_launch(g))

    def _residual(self, a0, a1, b0, b1):
        r = ctypes.c_double()
        check(self.L.lbm_residual(self.ctx.h, 2, a0.ptr, a1.ptr, None, b0.ptr, b1.ptr, None, self.N, ctypes.byref(r)))
        return r.value

    def _set_prec(self, prec):
        if self.pf.prec == prec:
            return
        for lat in (self.pf, self.pg):
            check(self.L.lbm_lattice_set_precision(lat.h, prec))
            lat.prec = prec
        for g in self._graphs.values():  # captured with the old buffers
            self.L.lbm_graph_destroy(g)
        self._graphs = {}

    def _solve(self, key, step, res):
        """Chunks of check_every steps until both residuals are below eps.  In FP32 / FP16 the
        residual is watched (switch_ready): once it is below fp32_floor, or has stalled below
        100 fp32_floor (ten checks) or STALL_CEILING (twenty checks), the lattices convert to FP64
        and the solve continues there.  Returns (steps, steps in reduced precision, final residuals)."""
        ce = self.check_every
        t, t32, best, stall = 0, 0, np.inf, 0
        r = (np.inf, np.inf)
        while t < self.max_steps:
            if t == 0:
                for s in range(1, ce + 1):
                    step(s)
            else:
                self._chunk(key, t, step)
            t += ce
            r = res(t)
            worst = max(r)
            if worst < self.eps or not np.isfinite(worst):  # converged, or diverged: stop either way
                break
            if self.pf.prec != F64:
                t32 = t
                stall = 0 if worst < 0.95 * best else stall + 1
                best = min(best, worst)
                if switch_ready(worst, self.fp32_floor, stall, self.stall_ceiling):
                    self._set_prec(F64)
        return t, (t32 if self.prec != F64 else 0), r

    def init_forward(self, warm=False):
        """Cold start of the forward solve (rest state, T = 0), or with warm=True (after
        keep_solution) the equilibrium of the kept converged fields."""
        L, d = self.L, self.d
        self._set_prec(self.prec)
        rho = d["one"]
        if warm and self.kept is not None:
            k = self.kept
            self._copy(d["ux0"], k["ux"])
            self._copy(d["uy0"], k["uy"])
            self._copy(d["tem"], k["tem"])
            rho = k["rho"]
        else:
            for name in ("ux0", "uy0"):
                d[name].fill(0.0)
            d["tem"].fill(0.0)
        check(L.lbm_ns_initial_condition(self.pf.h, rho.ptr, d["ux0"].ptr, d["uy0"].ptr, None))
        check(L.lbm_ad_initial_condition(self.pg.h, d["tem"].ptr, d["ux0"].ptr, d["uy0"].ptr, None))

    def residual_fields(self, adjoint, t):
        """(a, b) field pairs of the two residuals after step t: u and q (or their adjoints)."""
        d = self.d
        b, pb = t % 2, (t - 1) % 2
        u, q = (("iux", "iuy"), ("iqx", "iqy")) if adjoint else (("ux", "uy"), ("qx", "qy"))
        return ((d["%s%d" % (u[0], b)], d["%s%d" % (u[1], b)], d["%s%d" % (u[0], pb)], d["%s%d" % (u[1], pb)]),
                (d["%s%d" % (q[0], b)], d["%s%d" % (q[1], b)], d["%s%d" % (q[0], pb)], d["%s%d" % (q[1], pb)]))

    def forward(self, warm=False):
        self.init_forward(warm)
        res = lambda t: tuple(self._residual(*f) for f in self.residual_fields(False, t))
        t, self.fwd_steps32, self.fwd_res = self._solve("fwd", self._fwd_step, res)
        self.fwd_steps = t
        self.fwd_buf = t % 2
        return t

    def init_adjoint(self, warm=False):
        """Adjoint program for the converged forward velocity, and the adjoint cold start (or with
        warm=True, after keep_solution, the equilibrium of the kept adjoint fields)."""
        L, d = self.L, self.d
        self._set_prec(self.prec)
        fb = self.fwd_buf
        if self.adj_prog is None or getattr(self, "_adj_fb", None) != fb:
            if self.adj_prog is not None:
                L.lbm_program_destroy(self.adj_prog)
            self.adj_prog = self._record_adjoint(fb)
            self._adj_fb = fb
            g = self._graphs.pop("adj", None)
            if g is not None:
                L.lbm_graph_destroy(g)
        if warm and self.kept is not None:
            for name, src in (("ip", "ip"), ("item", "item"), ("iux0", "iux"), ("iuy0", "iuy"), ("iqx0", "iqx"),
                              ("iqy0", "iqy")):
                self._copy(d[name], self.kept[src])
        else:
            for k in ("ip", "item"):
                d[k].fill(0.0)
            for k in ("iux0", "iuy0", "iqx0", "iqy0"):
                d[k].fill(0.0)
        check(L.lbm_ans_initial_condition(self.pf.h, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, None, d["ip"].ptr,
                                          d["iux0"].ptr, d["iuy0"].ptr, None))
        check(L.lbm_aad_initial_condition(self.pg.h, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, None, d["item"].ptr,
                                          d["iqx0"].ptr, d["iqy0"].ptr, None))

    def adjoint(self, warm=False):
        self.init_adjoint(warm)
        fb = self.fwd_buf
        res = lambda t: tuple(self._residual(*f) for f in self.residual_fields(True, t))
        t, self.adj_steps32, self.adj_res = self._solve("adj", lambda s: self._adj_step(s, fb), res)
        self.adj_steps = t
        self.adj_buf = t % 2
        return t

    def sensitivity(self):
        """dF/dgamma: Brinkman and diffusivity terms (adjointadvection.h:1461-1475) plus the design
        derivative of the volumetric source Q(1 - gamma) (lbm_sensitivity_heat_source)."""
        L, d = self.L, self.d
        fb, ab = self.fwd_buf, self.adj_buf
        d["dfds"].fill(0.0)
        check(L.lbm_sensitivity_brinkman_diffusivity(
            self.pg.h, d["dfds"].ptr, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, None, d["imx"].ptr, d["imy"].ptr, None,
            d["dads"].ptr, d["tem"].ptr, d["item"].ptr, d["iqx%d" % ab].ptr, d["iqy%d" % ab].ptr, None, d["g"].ptr,
            d["ig"].ptr, d["diff"].ptr, d["dkds"].ptr))
        check(L.lbm_sensitivity_heat_source(
            self.pg.h, d["dfds"].ptr, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, None, d["item"].ptr,
            d["iqx%d" % ab].ptr, d["iqy%d" % ab].ptr, None, d["diff"].ptr, d["dhds"].ptr))
        out = ctypes.c_double()
        check(L.lbm_gather_sum(self.ctx.h, d["tem"].ptr, hptr(self.obj_nodes), len(self.obj_nodes), ctypes.byref(out)))
        self.objective = out.value / len(self.obj_nodes)
        return self.objective

    def in_fp64(self, solve, *args):
        """Run solve(*args) from its cold start in FP64, whatever this driver's storage precision."""
        prec = self.prec
        self.prec = F64
        try:
            return solve(*args)
        finally:
            self.prec = prec

    def evaluate(self, warm=False):
        """One design evaluation; returns wall seconds and step counts.  warm=True starts both
        solves from the last converged solution kept (keep_solution, done after every converged
        evaluation); a warm-started solve whose residual turns non-finite is repeated from the cold
        start (fallback).  A solve in reduced precision whose residual still turns non-finite is
        solved again from its cold start in FP64 (fp64_redo): a marginally unstable adjoint mode
        that FP64 round-off seeds too weakly to grow before the solve converges can grow from FP32
        round-off."""
        t0 = time.perf_counter()
        use = bool(warm) and self.kept is not None
        redo, fallback = [False, False], [False, False]
        self.forward(use)
        if use and not np.all(np.isfinite(self.fwd_res)):
            fallback[0] = True
            self.forward()
        if self.prec != F64 and not np.all(np.isfinite(self.fwd_res)):
            redo[0] = True
            self.in_fp64(self.forward)
        self.adjoint(use)
        if use and not np.all(np.isfinite(self.adj_res)):
            fallback[1] = True
            self.adjoint()
        if self.prec != F64 and not np.all(np.isfinite(self.adj_res)):
            redo[1] = True
            self.in_fp64(self.adjoint)
        self.sensitivity()
        res = np.concatenate([np.ravel(self.fwd_res), np.ravel(self.adj_res)])
        converged = bool(np.all(np.isfinite(res)) and res.max() < self.eps)
        if converged:
            self.keep_solution()
        self.ctx.sync()
        dt = time.perf_counter() - t0
        return {"seconds": dt, "fwd_steps": self.fwd_steps, "adj_steps": self.adj_steps,
                "fp32_steps": (self.fwd_steps32, self.adj_steps32), "objective": self.objective,
                "fwd_residual": self.fwd_res, "adj_residual": self.adj_res, "diverged": not np.all(np.isfinite(res)),
                "converged": converged, "warm": use, "fallback": fallback, "fp64_redo": redo}

    def close(self):
        for g in self._graphs.values():
            self.L.lbm_graph_destroy(g)
        self._graphs = {}
        for p in (self.fwd_prog, self.adj_prog):
            if p is not None:
                self.L.lbm_program_destroy(p)
        self.fwd_prog = self.adj_prog = None
