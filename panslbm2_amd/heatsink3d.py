"""3D heat-sink design evaluation on the HIP path: D3Q15 Navier-Stokes with Brinkman drag coupled to
D3Q7 advection-diffusion, forward steady solve, adjoint steady solve, design sensitivity.

The step order is the 2D driver's (panslbm2_amd/heatsink2d.py, after production/heatsink.cpp:
136-232) on the 3D overloads (advection.h:871, adjointadvection.h:1054, :1478), with the boundary
set of a duct heat sink:
  * inlet x = 0 (face interior): uniform velocity u_in (Zou-He), T = 0;
  * outlet x = lx-1 (face interior): pressure rho = 1 (Zou-He), zero heat flux;
  * walls y = 0, ly-1, z = 0, lz-1: bounce-back, zero heat flux;
  * base plate z < plate: solid (gamma = 0) with a volumetric heat source Q;
  * design gamma elsewhere: seeded uniform[0,1] smoothed by the density filter (radius 2);
  * objective: mean temperature on the bottom face z = 0, injected into the adjoint through the flux
    condition with eps = 1 (heatsink.cpp:202); the adjoint outlet is ANS::iBoundaryConditionSetRho3D
    (the reference's 3D AAD::iBoundaryConditionSetRho does not compile).
By default no SmoothCorner: the corner and edge nodes lie on bounce-back walls, and each time step
is one fused kernel pass.  With smooth_corners=True both lattices are smoothed after the boundary
calls of every forward and adjoint step, as the reference's 3D driver does (heatsink3D.cpp:175-176,
215-216); the fused pass then skips the edge and corner nodes, which a kernel on the context's
second stream smooths, collides and stores concurrently with the pass (107.8 vs 91.7 us per step at
128x128x64).
The Navier-Stokes lattice is D3Q15 (the reference's) or D3Q27 (ns_kind=D3Q27, no reference
counterpart; its velocity and pressure boundaries are non-equilibrium bounce-back).
"""
import ctypes
import time

import numpy as np

from . import D3Q7, D3Q15, D3Q27, F16, F32, F64, NC, Context, Lattice, check, hptr, to_face
from .heatsink2d import Heatsink2D

D = ctypes.c_double
VP = ctypes.c_void_p


class Heatsink3D(Heatsink2D):
    # no plateau clause above 100x the switch floor: the 3D duct's FP32 residual converges through
    # 1e-5, and its transient plateau between 1e-5 and 1e-4 would be taken for the FP32 floor
    # (62 % instead of 81 % of the 128x128x64 evaluation's steps in FP32)
    stall_ceiling = 0.0

    def __init__(self, ctx=None, lx=128, ly=128, lz=64, ad_kind=D3Q7, seed=0, nu=0.1, pr=6.0, uin=0.03, heat=1e-3,
                 plate=8, alpha_max=1e4, qf=0.01, qg=1.0, filter_radius=2.0, eps=1e-6, check_every=100,
                 max_steps=400000, use_graphs=True, prec=F64, fp32_floor=None, fused=True, ns_kind=D3Q15,
                 heaviside_beta=None, design_box=None, smooth_corners=False, inlet_rho=None):
        assert ns_kind in (D3Q15, D3Q27)
        self.ctx = ctx or Context(0)
        self.L = self.ctx.lib
        self.lx, self.ly, self.lz = lx, ly, lz
        self.N = N = lx * ly * lz
        self.ad_kind = ad_kind
        self.nu, self.eps, self.check_every, self.max_steps = nu, eps, check_every, max_steps
        assert check_every % 2 == 0
        self.use_graphs, self.fused = use_graphs and fused, fused
        self.smooth_corners = smooth_corners
        # inlet_rho: a pressure inlet (Zou-He density rho_in, navierstokes.h:261-426, and its adjoint
        # ANS::iBoundaryConditionSetRho3D) instead of the velocity inlet: the flow rate then follows
        # the design, so a design that narrows the duct slows the flow instead of accelerating it
        self.inlet_rho = inlet_rho
        self.prec = prec
        self.fp32_floor = fp32_floor if fp32_floor is not None else (1e-3 if prec == F16 else 1e-7)
        c = self.ctx
        n = np.arange(N)
        i, j, k = n % lx, (n // lx) % ly, n // (lx * ly)
        self.i, self.j, self.k = i, j, k

        self.plate, self.filter_radius = plate, filter_radius
        self.heaviside_beta = heaviside_beta
        # the Heaviside filter's design region [0,mx) x [0,my) x [0,mz) (heatsink3D.cpp:87-93); the
        # whole grid by default (the filter's default weights)
        self.design_box = tuple(design_box) if design_box is not None else (lx, ly, lz)
        # as the reference's driver: nodes outside the box are fluid (heatsink3D.cpp:104-108, 286-290),
        # and the design variables are the box's nodes above the plate
        mx, my, mz = self.design_box
        self.in_box = (i < mx) & (j < my) & (k < mz)
        self.region = self.in_box & (k >= plate)
        self.alpha_max, self.qf, self.qg = alpha_max, qf, qg
        self.dfl = nu / pr
        self.dso = 10.0 * self.dfl
        rng = np.random.default_rng(seed)
        self.d = {name: c.array((N,), fill=0.0)
                  for name in ("alpha", "dads", "diff", "dkds", "raw", "design", "dgamma")}
        self.d["heat"] = c.array((N,), data=np.where(k < plate, heat, 0.0))
        self.set_design(rng.uniform(0.0, 1.0, N))
        for name in ("rho", "tem", "ip", "imx", "imy", "imz", "item", "dfds"):
            self.d[name] = c.array((N,), fill=0.0)
        for name in ("ux", "uy", "uz", "qx", "qy", "qz", "iux", "iuy", "iuz", "iqx", "iqy", "iqz"):
            self.d[name + "0"] = c.array((N,), fill=0.0)
            self.d[name + "1"] = c.array((N,), fill=0.0)
        self.d["g"] = c.array((NC[ad_kind], N), fill=0.0)
        self.d["ig"] = c.array((NC[ad_kind], N), fill=0.0)
        self.d["one"] = c.array((N,), fill=1.0)

        self.ns_kind = ns_kind
        self.pf = Lattice(c, ns_kind, lx, ly, lz, prec=prec)
        self.pg = Lattice(c, ad_kind, lx, ly, lz, prec=prec)

        wall = (j == 0) | (j == ly - 1) | (k == 0) | (k == lz - 1)
        inner = ~wall
        inlet = ((i == 0) & inner).astype(np.int32)
        outlet = ((i == lx - 1) & inner).astype(np.int32)
        flux = (((i > 0) | wall)).astype(np.int32)  # every boundary node but the inlet's interior
        bottom = (k == 0).astype(np.int32)
        self.obj_nodes = np.nonzero(k == 0)[0].astype(np.int32)
        F = lambda kind, a, dt=np.float64: to_face(kind, lx, ly, lz, a, dt)
        zero = np.zeros(N)
        self.faces = dict(
            walls_f=F(D3Q15, wall.astype(np.int32), np.int32), inlet_f=F(D3Q15, inlet, np.int32),
            outlet_f=F(D3Q15, outlet, np.int32), uin_f=F(D3Q15, np.full(N, uin)), zero_f=F(D3Q15, zero),
            rhoin_f=F(D3Q15, np.full(N, inlet_rho if inlet_rho is not None else 1.0)),
            one_f=F(D3Q15, np.ones(N)), inlet_g=F(ad_kind, inlet, np.int32), flux_g=F(ad_kind, flux, np.int32),
            obj_g=F(ad_kind, bottom, np.int32), zero_g=F(ad_kind, zero),
        )
        self._graphs = {}
        self.adj_prog = None
        self.fwd_prog = self._record_forward() if fused else None

    # ---- design -----------------------------------------------------------------------------------
    def set_design(self, x):
        """Design variables x (fluid fraction before filtering, per node) -> gamma = density-filtered
        x (or, with heaviside_beta, its Heaviside projection, heatsink3D.cpp:103) with the base plate
        solid -> Brinkman and diffusivity coefficients (heatsink.cpp:104-117 interpolation; 3D duct
        geometry)."""
        c, lx, ly, lz, d = self.ctx, self.lx, self.ly, self.lz, self.d
        self.x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 1.0).copy()
        self.x[~self.in_box] = 1.0
        d["raw"].upload(self.x)
        if self.heaviside_beta:
            check(self.L.lbm_heaviside_filter_region(c.h, lx, ly, lz, D(self.filter_radius), D(self.heaviside_beta),
                                                     *self.design_box, d["raw"].ptr, d["design"].ptr))
        else:
            check(self.L.lbm_density_filter(c.h, lx, ly, lz, D(self.filter_radius), d["raw"].ptr,
                                            d["design"].ptr))
        s = d["design"].download()
        s[~self.in_box] = 1.0  # fluid outside the design box
        s[self.k < self.plate] = 0.0  # the base plate is solid
        d["design"].upload(s)
        prev = getattr(self, "gamma", None)
        d["dgamma"].upload(s - prev if prev is not None else np.zeros_like(s))
        self.gamma = s
        am, qf, qg, dfl, dso = self.alpha_max, self.qf, self.qg, self.dfl, self.dso
        d["alpha"].upload(am / (ly - 1.0) * qf * (1.0 - s) / (s + qf))
        d["dads"].upload(-am / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2)
        d["diff"].upload(dso + (dfl - dso) * s * (1.0 + qg) / (s + qg))
        d["dkds"].upload((dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2)

    def design_update(self, vf=0.4, move=0.05, optimizer="mma"):
        """One design step over the design variables (the design box's nodes above the plate).
        optimizer="mma": the device MMA
        optimiser (panslbm2_amd/optimizer.py) as the reference's driver uses it (heatsink3D.cpp:78-85,
        :120-139, :282): one constraint g = mean(1 - x) / vf - 1 <= 0 (solid fraction at most vf),
        a0 = 1, a = 0, c = 1e4, d = 0, the move limit, and the sensitivity normalised to a largest
        magnitude of 1 (normalize.h:9-24).  optimizer="pg": a projected-gradient step
        x <- clip(x - move dF/ds / max|dF/ds| + lam, x - move, x + move, 0, 1) with lam found by
        bisection so that the solid fraction equals vf.

        With heaviside_beta the gradients reach x through the projection's sensitivity (the
        reference driver's chain, heatsink3D.cpp:249-250), with dF/dgamma zero where gamma is set
        after the filter (base plate, outside the box).  Without it, the updates take dF/dgamma
        as the gradient in x: the density filter's own chain is not applied."""
        if optimizer == "mma":
            return self._mma_update(vf, move)
        dfds = self.d["dfds"].download()
        region = self.region
        g = dfds[region]
        g = g / max(np.abs(g).max(), 1e-300)
        x = self.x[region]
        lo_b, hi_b = np.maximum(0.0, x - move), np.minimum(1.0, x + move)
        trial = lambda lam: np.clip(x - move * g + lam, lo_b, hi_b)
        lo, hi = -1.0, 1.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            if trial(mid).mean() > 1.0 - vf:
                hi = mid
            else:
                lo = mid
        xn = self.x.copy()
        xn[region] = trial(0.5 * (lo + hi))
        self.set_design(xn)

    def _mma_update(self, vf, move):
        from .optimizer import MMA

        region = np.nonzero(self.region)[0]
        n = len(region)
        c = self.ctx
        if getattr(self, "_mma", None) is None:
            self._mma = MMA(c, n, 1, 1.0, [0.0], [1e4], [0.0], np.zeros(n), np.ones(n), move=move)
            self._mx = c.array((n,))
            self._mdf = c.array((n,))
            self._mdg = c.array((1, n), data=np.full(n, -1.0 / (vf * n)))
        x = self.x[region]
        if self.heaviside_beta:
            # through the filter (heatsink3D.cpp:249-250): the constraint on the projected design
            d, lx, ly, lz = self.d, self.lx, self.ly, self.lz
            R, beta = D(self.filter_radius), D(self.heaviside_beta)
            if "dfdx" not in d:
                d["dfdx"] = c.array((self.N,))
                d["dfdg"] = c.array((self.N,))
                d["dgdg"] = c.array((self.N,), data=np.where(self.region, -1.0 / (vf * n), 0.0))
                d["dgdx"] = c.array((self.N,))
            box = self.design_box
            # gamma of the base plate (and outside the box) is set after the filter, whatever x is:
            # its sensitivity must not flow back into the design variables near the plate
            d["dfdg"].upload(np.where(self.region, d["dfds"].download(), 0.0))
            check(self.L.lbm_heaviside_sensitivity_region(c.h, lx, ly, lz, R, beta, *box, d["raw"].ptr,
                                                          d["dfdg"].ptr, d["dfdx"].ptr))
            check(self.L.lbm_heaviside_sensitivity_region(c.h, lx, ly, lz, R, beta, *box, d["raw"].ptr,
                                                          d["dgdg"].ptr, d["dgdx"].ptr))
            dfds = d["dfdx"].download()[region]
            self._mdg.upload(d["dgdx"].download()[region][None])
            g = (1.0 - self.gamma[region]).sum() / (vf * n) - 1.0
        else:
            dfds = self.d["dfds"].download()[region]
            g = (1.0 - x).sum() / (vf * n) - 1.0
        dfds = dfds / max(np.abs(dfds).max(), 1e-300)
        self._mx.upload(x)
        self._mdf.upload(dfds)
        self.mma_iterations = self._mma.update(self._mx, self.objective, self._mdf, [g], self._mdg)
        xn = self.x.copy()
        xn[region] = self._mx.download()
        self.set_design(xn)

    # ---- boundary calls of one step (recorded into a program, or run eagerly) ----------------------
    def _fwd_boundaries(self, ux, uy, uz):
        L, d, fc = self.L, self.d, self.faces
        check(L.lbm_stream(self.pf.h))
        check(L.lbm_stream(self.pg.h))
        check(L.lbm_bc(self.pf.h, hptr(fc["walls_f"])))
        if self.inlet_rho is None:
            check(L.lbm_ns_set_u(self.pf.h, hptr(fc["inlet_f"]), hptr(fc["uin_f"]), hptr(fc["zero_f"]),
                                 hptr(fc["zero_f"])))
        else:
            check(L.lbm_ns_set_rho(self.pf.h, hptr(fc["inlet_f"]), hptr(fc["rhoin_f"]), hptr(fc["zero_f"]),
                                   hptr(fc["zero_f"])))
        check(L.lbm_ns_set_rho(self.pf.h, hptr(fc["outlet_f"]), hptr(fc["one_f"]), hptr(fc["zero_f"]),
                               hptr(fc["zero_f"])))
        check(L.lbm_ad_set_t(self.pg.h, hptr(fc["inlet_g"]), hptr(fc["zero_g"]), ux, uy, uz))
        check(L.lbm_ad_set_q(self.pg.h, hptr(fc["flux_g"]), hptr(fc["zero_g"]), ux, uy, uz, d["diff"].ptr, D(0.0)))
        if self.smooth_corners:
            check(L.lbm_smooth_corner(self.pf.h))
            check(L.lbm_smooth_corner(self.pg.h))

    def _adj_boundaries(self, ux, uy, uz):
        L, fc = self.L, self.faces
        check(L.lbm_istream(self.pf.h))
        check(L.lbm_istream(self.pg.h))
        check(L.lbm_aad_iset_t(self.pg.h, hptr(fc["inlet_g"]), ux, uy, uz))
        check(L.lbm_aad_iset_q(self.pg.h, hptr(fc["flux_g"]), ux, uy, uz, D(0.0)))
        check(L.lbm_aad_iset_q(self.pg.h, hptr(fc["obj_g"]), ux, uy, uz, D(1.0)))
        check(L.lbm_ibc(self.pf.h, hptr(fc["walls_f"])))
        if self.inlet_rho is None:
            check(L.lbm_ans_iset_u(self.pf.h, hptr(fc["inlet_f"]), hptr(fc["uin_f"]), hptr(fc["zero_f"]),
                                   hptr(fc["zero_f"]), D(0.0), 1))
        else:
            check(L.lbm_ans_iset_rho(self.pf.h, hptr(fc["inlet_f"])))
        check(L.lbm_ans_iset_rho(self.pf.h, hptr(fc["outlet_f"])))
        if self.smooth_corners:
            check(L.lbm_smooth_corner(self.pf.h))
            check(L.lbm_smooth_corner(self.pg.h))

    def _record(self, calls):
        check(self.L.lbm_record_begin(self.ctx.h))
        calls()
        prog = VP()
        check(self.L.lbm_record_end(self.ctx.h, ctypes.byref(prog)))
        # every program here fuses (with SmoothCorner too: the edges and corners run after the pass)
        assert self.L.lbm_program_fusable(prog) == 1
        return prog

    def _record_forward(self):
        d = self.d
        return self._record(lambda: self._fwd_boundaries(d["ux0"].ptr, d["uy0"].ptr, d["uz0"].ptr))

    def _record_adjoint(self, b):
        d = self.d
        return self._record(lambda: self._adj_boundaries(d["ux%d" % b].ptr, d["uy%d" % b].ptr, d["uz%d" % b].ptr))

    # ---- steps ------------------------------------------------------------------------------------
    def _fwd_step(self, t):
        L, d = self.L, self.d
        b, pb = t % 2, (t - 1) % 2
        prog = None
        if t > 1:
            if self.fused:
                prog = self.fwd_prog
                for s in range(L.lbm_program_num_field_sets(prog)):
                    check(L.lbm_program_bind_fields(prog, s, None, d["ux%d" % pb].ptr, d["uy%d" % pb].ptr,
                                                    d["uz%d" % pb].ptr, None, None))
            else:
                self._fwd_boundaries(d["ux%d" % pb].ptr, d["uy%d" % pb].ptr, d["uz%d" % pb].ptr)
        mode = self._save_mode(t, False)
        check(L.lbm_ad_macro_brinkman_collide_force_convection_p(
            prog, self.pf.h, d["rho"].ptr, d["ux%d" % b].ptr, d["uy%d" % b].ptr, d["uz%d" % b].ptr, d["alpha"].ptr,
            D(self.nu), self.pg.h, d["tem"].ptr, d["qx%d" % b].ptr, d["qy%d" % b].ptr, d["qz%d" % b].ptr,
            d["diff"].ptr, mode, d["g"].ptr, d["heat"].ptr))

    def _adj_step(self, t, fb):
        L, d = self.L, self.d
        b = t % 2
        prog = None
        if t > 1:
            if self.fused:
                prog = self.adj_prog
            else:
                self._adj_boundaries(d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, d["uz%d" % fb].ptr)
        mode = self._save_mode(t, True)
        check(L.lbm_aad_macro_brinkman_collide_force_convection_p(
            prog, self.pf.h, d["rho"].ptr, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, d["uz%d" % fb].ptr, d["ip"].ptr,
            d["iux%d" % b].ptr, d["iuy%d" % b].ptr, d["iuz%d" % b].ptr, d["imx"].ptr, d["imy"].ptr, d["imz"].ptr,
            d["alpha"].ptr, D(self.nu), self.pg.h, d["tem"].ptr, d["item"].ptr, d["iqx%d" % b].ptr,
            d["iqy%d" % b].ptr, d["iqz%d" % b].ptr, d["diff"].ptr, mode, d["ig"].ptr))

    def _residual3(self, a, b):
        r = ctypes.c_double()
        check(self.L.lbm_residual(self.ctx.h, 3, a[0].ptr, a[1].ptr, a[2].ptr, b[0].ptr, b[1].ptr, b[2].ptr, self.N,
                                  ctypes.byref(r)))
        return r.value

    def residual_fields(self, adjoint, t):
        d = self.d
        b, pb = t % 2, (t - 1) % 2
        u, q = (("iux", "iuy", "iuz"), ("iqx", "iqy", "iqz")) if adjoint else (("ux", "uy", "uz"), ("qx", "qy", "qz"))
        return tuple(([d["%s%d" % (v, b)] for v in f], [d["%s%d" % (v, pb)] for v in f]) for f in (u, q))

    def init_forward(self, warm=False):
        """Cold start (rest, T = 0), or with warm=True the equilibrium of the predicted fields."""
        L, d = self.L, self.d
        self._set_prec(self.prec)
        if warm:
            w = self.w
            check(L.lbm_ns_initial_condition(self.pf.h, w["rho"].ptr, w["ux"].ptr, w["uy"].ptr, w["uz"].ptr))
            check(L.lbm_ad_initial_condition(self.pg.h, w["tem"].ptr, w["ux"].ptr, w["uy"].ptr, w["uz"].ptr))
            return
        for name in ("ux0", "uy0", "uz0"):
            d[name].fill(0.0)
        d["tem"].fill(0.0)
        check(L.lbm_ns_initial_condition(self.pf.h, d["one"].ptr, d["ux0"].ptr, d["uy0"].ptr, d["uz0"].ptr))
        check(L.lbm_ad_initial_condition(self.pg.h, d["tem"].ptr, d["ux0"].ptr, d["uy0"].ptr, d["uz0"].ptr))

    def forward(self, warm=False):
        self.init_forward(warm)
        res = lambda t: tuple(self._residual3(*f) for f in self.residual_fields(False, t))
        t, self.fwd_steps32, self.fwd_res = self._solve("fwd", self._fwd_step, res)
        self.fwd_steps = t
        self.fwd_buf = t % 2
        return t

    def init_adjoint(self, warm=False):
        L, d = self.L, self.d
        self._set_prec(self.prec)
        fb = self.fwd_buf
        if self.fused and (self.adj_prog is None or getattr(self, "_adj_fb", None) != fb):
            if self.adj_prog is not None:
                L.lbm_program_destroy(self.adj_prog)
            self.adj_prog = self._record_adjoint(fb)
            self._adj_fb = fb
            g = self._graphs.pop("adj", None)
            if g is not None:
                L.lbm_graph_destroy(g)
        uf = (d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, d["uz%d" % fb].ptr)
        if warm:
            w = self.w
            check(L.lbm_ans_initial_condition(self.pf.h, *uf, w["ip"].ptr, w["iux"].ptr, w["iuy"].ptr, w["iuz"].ptr))
            check(L.lbm_aad_initial_condition(self.pg.h, *uf, w["item"].ptr, w["iqx"].ptr, w["iqy"].ptr,
                                              w["iqz"].ptr))
            return
        for name in ("ip", "item"):
            d[name].fill(0.0)
        for name in ("iux0", "iuy0", "iuz0", "iqx0", "iqy0", "iqz0"):
            d[name].fill(0.0)
        check(L.lbm_ans_initial_condition(self.pf.h, *uf, d["ip"].ptr, d["iux0"].ptr, d["iuy0"].ptr, d["iuz0"].ptr))
        check(L.lbm_aad_initial_condition(self.pg.h, *uf, d["item"].ptr, d["iqx0"].ptr, d["iqy0"].ptr,
                                          d["iqz0"].ptr))

    def adjoint(self, warm=False):
        self.init_adjoint(warm)
        fb = self.fwd_buf
        res = lambda t: tuple(self._residual3(*f) for f in self.residual_fields(True, t))
        t, self.adj_steps32, self.adj_res = self._solve("adj", lambda s: self._adj_step(s, fb), res)
        self.adj_steps = t
        self.adj_buf = t % 2
        return t

    def sensitivity(self):
        """dF/dgamma: Brinkman and diffusivity terms (adjointadvection.h:1478-1493)."""
        L, d = self.L, self.d
        fb, ab = self.fwd_buf, self.adj_buf
        d["dfds"].fill(0.0)
        check(L.lbm_sensitivity_brinkman_diffusivity(
            self.pg.h, d["dfds"].ptr, d["ux%d" % fb].ptr, d["uy%d" % fb].ptr, d["uz%d" % fb].ptr, d["imx"].ptr,
            d["imy"].ptr, d["imz"].ptr, d["dads"].ptr, d["tem"].ptr, d["item"].ptr, d["iqx%d" % ab].ptr,
            d["iqy%d" % ab].ptr, d["iqz%d" % ab].ptr, d["g"].ptr, d["ig"].ptr, d["diff"].ptr, d["dkds"].ptr))
        out = ctypes.c_double()
        check(L.lbm_gather_sum(self.ctx.h, d["tem"].ptr, hptr(self.obj_nodes), len(self.obj_nodes), ctypes.byref(out)))
        self.objective = out.value / len(self.obj_nodes)
        return self.objective

    # ---- learned warm start -----------------------------------------------------------------------
    WARM_OUT = ("rho", "ux", "uy", "uz", "tem", "ip", "iux", "iuy", "iuz", "item", "iqx", "iqy", "iqz")

    def attach_warm_start(self, seed=0, lr=1e-3, steps=40, replay=4, learn=True):
        """A WarmStart network for this grid.  Inputs: gamma, its change since the previous design,
        x/(lx-1), z/(lz-1) (the coordinates tell inlet from outlet and plate from lid); outputs:
        the change of the fields of WARM_OUT, which the forward (NS, AD) and adjoint (ANS, AAD)
        InitialCondition calls take, since the previous converged design."""
        from .warmstart import WarmStart

        c = self.ctx
        self.d["xpos"] = c.array((self.N,), data=self.i / (self.lx - 1.0))
        self.d["zpos"] = c.array((self.N,), data=self.k / (self.lz - 1.0))
        self.w = {n: c.array((self.N,), fill=0.0) for n in self.WARM_OUT}
        self.warm = WarmStart(c, (self.lx, self.ly, self.lz), 4, len(self.WARM_OUT), seed=seed, lr=lr, steps=steps,
                              replay=replay, learn=learn)
        return self.warm

    def _warm_inputs(self):
        return [self.d["design"], self.d["dgamma"], self.d["xpos"], self.d["zpos"]]

    def _warm_targets(self):
        d, fb, ab = self.d, self.fwd_buf, self.adj_buf
        return [d["rho"], d["ux%d" % fb], d["uy%d" % fb], d["uz%d" % fb], d["tem"], d["ip"], d["iux%d" % ab],
                d["iuy%d" % ab], d["iuz%d" % ab], d["item"], d["iqx%d" % ab], d["iqy%d" % ab], d["iqz%d" % ab]]

    def evaluate(self, warm=None, train=True):
        """One design evaluation: forward solve, adjoint solve, sensitivity.

        warm=None starts both solves from the attached warm start's prediction once it has seen a
        converged solution (warm=False forces the cold start); a warm-started solve whose residual
        turns non-finite is repeated from the cold start, and a reduced-precision one that still
        does, from the cold start in FP64.  With train=True the warm start's
        training on the stored solutions is queued right after its prediction and runs beside the
        solves, and a converged evaluation is then stored as a new solution.  'seconds' is the
        wall time of the solves (training beside them included), 'train_seconds' the time to
        store the solution afterwards."""
        ws = getattr(self, "warm", None)
        use = ws is not None and ws.ready and warm is not False
        t0 = time.perf_counter()
        fallback = [False, False]
        if use:
            ws.predict(self._warm_inputs(), [self.w[n] for n in self.WARM_OUT])
        if ws is not None and train:
            ws.train()
        redo = [False, False]
        self.forward(use)
        if use and not np.all(np.isfinite(self.fwd_res)):
            fallback[0] = True
            self.forward(False)
        if self.prec != F64 and not np.all(np.isfinite(self.fwd_res)):
            redo[0] = True
            self.in_fp64(self.forward, False)
        self.adjoint(use)
        if use and not np.all(np.isfinite(self.adj_res)):
            fallback[1] = True
            self.adjoint(False)
        if self.prec != F64 and not np.all(np.isfinite(self.adj_res)):
            redo[1] = True
            self.in_fp64(self.adjoint, False)
        self.sensitivity()
        self.ctx.sync()
        t1 = time.perf_counter()
        res = np.concatenate([np.ravel(self.fwd_res), np.ravel(self.adj_res)])
        converged = bool(np.all(np.isfinite(res)) and res.max() < self.eps)
        if ws is not None and train and converged:
            ws.observe(self._warm_inputs(), self._warm_targets())
            self.ctx.sync()
        t2 = time.perf_counter()
        return {"seconds": t1 - t0, "train_seconds": t2 - t1, "fwd_steps": self.fwd_steps,
                "adj_steps": self.adj_steps, "fp32_steps": (self.fwd_steps32, self.adj_steps32),
                "objective": self.objective, "fwd_residual": self.fwd_res, "adj_residual": self.adj_res,
                "diverged": not np.all(np.isfinite(res)), "converged": converged, "warm": use,
                "fallback": fallback, "fp64_redo": redo, "warm_gain": None if ws is None else ws.gain.tolist()}

    def close(self):
        super().close()
        if getattr(self, "_mma", None) is not None:
            self._mma.close()
            self._mma = None
        ws = getattr(self, "warm", None)
        if ws is not None:
            ws.close()
            self.warm = None

    # the reference driver's continuation (heatsink3D.cpp:45,100-118,402-405): (qf, beta) per stage
    REFERENCE_STAGES = ((1e-2, 1.0), (1e-1, 2.0), (1e-1, 4.0), (1e-1, 8.0), (1e-1, 16.0))

    def optimize(self, iterations, vf=0.4, move=0.05, warm=None, x0=None, callback=None, optimizer="mma",
                 stages=None, stage_iterations=100, stage_tol=0.01):
        """Design loop from a uniform solid fraction vf over the design variables (the design box's
        nodes above the plate), or from x0: evaluate, then design_update.  Returns the per-iteration
        evaluate() records.

        stages: a continuation schedule of (qf, beta) pairs as the reference's driver runs it
        (REFERENCE_STAGES): stage s sets the Brinkman interpolation parameter qf and, with a
        Heaviside projection, its sharpness beta; the loop moves to the next stage when the largest
        design change of an update falls below stage_tol or after stage_iterations iterations in
        the stage, and stops in the last stage once the change is below stage_tol."""
        if stages:
            self.qf = stages[0][0]
            if self.heaviside_beta:
                self.heaviside_beta = stages[0][1]
        x = np.where(self.region, 1.0 - vf, 0.0) if x0 is None else x0
        self.set_design(x)
        old = getattr(self, "_mma", None)
        if old is not None:
            old.close()
        self._mma = None
        hist = []
        stage, cnt = 0, 0
        for it in range(iterations):
            r = self.evaluate(warm=warm)
            r["iteration"] = it
            r["stage"] = stage
            hist.append(r)
            if callback is not None:
                callback(r)
            if it + 1 < iterations:
                x_old = self.x.copy()
                self.design_update(vf, move, optimizer)
                if stages:
                    cnt += 1
                    dsmax = float(np.abs(self.x - x_old).max())
                    r["dsmax"] = dsmax
                    if stage == len(stages) - 1 and dsmax < stage_tol:
                        break
                    if dsmax < stage_tol or cnt % stage_iterations == 0:
                        stage, cnt = min(stage + 1, len(stages) - 1), 0
                        self.qf = stages[stage][0]
                        if self.heaviside_beta:
                            self.heaviside_beta = stages[stage][1]
                        self.set_design(self.x)  # the new stage's coefficients for the same design
        ws = getattr(self, "warm", None)
        if ws is not None:
            ws.net.synchronize()
        return hist
