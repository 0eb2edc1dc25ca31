"""The driver-sequence backend (tests/sequences.py) on the HIP library, through its C ABI.

GpuBackend runs every reference call as its own C-ABI call (eager).  FusedGpuBackend records the
stream + boundary + SmoothCorner calls of a step and hands the program to the next collide
(*_collide_p), i.e. the fused single-pass kernels.
"""
import ctypes

import numpy as np

import panslbm2_amd as pa
from panslbm2_amd import D, check, hptr

VP = ctypes.c_void_p


class GpuBackend:
    name = "gpu"

    def __init__(self, ctx=None, prec=pa.F64):
        self.ctx = ctx or pa.Context(0)
        self.L = self.ctx.lib
        self.fields = {}
        self.prec = prec

    # ---- fields -------------------------------------------------------------------------
    def put(self, name, arr):
        arr = np.asarray(arr, dtype=np.float64)
        self.fields[name] = self.ctx.array(arr.shape, data=arr)

    def zeros(self, name, n):
        shape = n if isinstance(n, tuple) else (n,)
        self.fields[name] = self.ctx.array(shape, fill=0.0)

    def get(self, name):
        return self.fields[name].download()

    def swap(self, a, b):
        self.fields[a], self.fields[b] = self.fields[b], self.fields[a]

    def f(self, name):
        return None if name is None else self.fields[name].ptr

    # ---- lattices -----------------------------------------------------------------------
    def lattice(self, kind, lx, ly, lz=1):
        return pa.Lattice(self.ctx, kind, lx, ly, lz, prec=self.prec)

    def pops(self, h):
        return h.download()

    def set_pops(self, h, soa):
        h.upload(soa)

    def _face(self, h, a, dt=np.float64):
        return pa.to_face(h.kind, h.lx, h.ly, h.lz, a, dt)

    def _fi(self, h, a):
        return self._face(h, a, np.int32)

    def stream(self, h):
        check(self.L.lbm_stream(h.h))

    def istream(self, h):
        check(self.L.lbm_istream(h.h))

    def bc(self, h, type_node, adjoint=False):
        fn = self.L.lbm_ibc if adjoint else self.L.lbm_bc
        check(fn(h.h, hptr(self._fi(h, type_node))))

    def smooth(self, h):
        check(self.L.lbm_smooth_corner(h.h))

    def ns_set_u(self, h, mask, ux, uy, uz=None):
        uz = np.zeros_like(ux) if uz is None else uz
        check(self.L.lbm_ns_set_u(h.h, hptr(self._fi(h, mask)), hptr(self._face(h, ux)), hptr(self._face(h, uy)),
                                  hptr(self._face(h, uz))))

    def ns_set_rho(self, h, mask, rho, us, ut=None):
        ut = np.zeros_like(us) if ut is None else ut
        check(self.L.lbm_ns_set_rho(h.h, hptr(self._fi(h, mask)), hptr(self._face(h, rho)), hptr(self._face(h, us)),
                                    hptr(self._face(h, ut))))

    def ad_set_t(self, h, mask, tem, ux, uy, uz=None):
        check(self.L.lbm_ad_set_t(h.h, hptr(self._fi(h, mask)), hptr(self._face(h, tem)), self.f(ux), self.f(uy),
                                  self.f(uz)))

    def ad_set_q(self, h, mask, qn, ux, uy, uz, diff):
        field = isinstance(diff, str)
        check(self.L.lbm_ad_set_q(h.h, hptr(self._fi(h, mask)), hptr(self._face(h, qn)), self.f(ux), self.f(uy),
                                  self.f(uz), self.f(diff) if field else None, D(0.0 if field else float(diff))))

    def ans_iset_u(self, h, mask, ux, uy, uz=None, eps=0.0):
        uz = np.zeros_like(ux) if uz is None else uz
        check(self.L.lbm_ans_iset_u(h.h, hptr(self._fi(h, mask)), hptr(self._face(h, ux)), hptr(self._face(h, uy)),
                                    hptr(self._face(h, uz)), D(eps), 1))

    def ans_iset_rho(self, h, mask):
        check(self.L.lbm_ans_iset_rho(h.h, hptr(self._fi(h, mask))))

    def aad_iset_t(self, h, mask, ux, uy, uz=None):
        check(self.L.lbm_aad_iset_t(h.h, hptr(self._fi(h, mask)), self.f(ux), self.f(uy), self.f(uz)))

    def aad_iset_q(self, h, mask, ux, uy, uz=None, eps=0.0):
        check(self.L.lbm_aad_iset_q(h.h, hptr(self._fi(h, mask)), self.f(ux), self.f(uy), self.f(uz), D(eps)))

    def aad_iset_rho(self, p, q, type_node, rho, ux, uy, uz, tem, eps=0.0):
        check(self.L.lbm_aad_iset_rho(p.h, q.h, hptr(self._fi(p, type_node)), self.f(rho), self.f(ux), self.f(uy),
                                      self.f(uz), self.f(tem), D(eps)))

    def ns_init(self, h, rho, ux, uy, uz=None):
        check(self.L.lbm_ns_initial_condition(h.h, self.f(rho), self.f(ux), self.f(uy), self.f(uz)))

    def ad_init(self, h, tem, ux, uy, uz=None):
        check(self.L.lbm_ad_initial_condition(h.h, self.f(tem), self.f(ux), self.f(uy), self.f(uz)))

    def ans_init(self, h, ux, uy, uz, ip, iux, iuy, iuz):
        check(self.L.lbm_ans_initial_condition(h.h, *[self.f(n) for n in (ux, uy, uz, ip, iux, iuy, iuz)]))

    def aad_init(self, h, ux, uy, uz, item, iqx, iqy, iqz):
        check(self.L.lbm_aad_initial_condition(h.h, *[self.f(n) for n in (ux, uy, uz, item, iqx, iqy, iqz)]))

    # collides: the eager backend passes no program
    def _prog(self):
        return None

    def ns_collide(self, h, rho, ux, uy, uz, nu, save=True):
        check(self.L.lbm_ns_macro_collide_p(self._prog(), h.h, *[self.f(n) for n in (rho, ux, uy, uz)], D(nu),
                                            int(save)))

    def ns_brinkman_collide(self, h, rho, ux, uy, uz, nu, alpha, save=True):
        check(self.L.lbm_ns_macro_brinkman_collide_p(self._prog(), h.h, *[self.f(n) for n in (rho, ux, uy, uz)],
                                                     D(nu), self.f(alpha), int(save)))

    def ad_collide_fc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, save=True, gsnap=None,
                      heat=None):
        check(self.L.lbm_ad_macro_brinkman_collide_force_convection_p(
            self._prog(), p.h, *[self.f(n) for n in (rho, ux, uy, uz, alpha)], D(nu), q.h,
            *[self.f(n) for n in (tem, qx, qy, qz, diff)], int(save), self.f(gsnap), self.f(heat)))

    def ad_collide_nc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, gx, gy, gz, tem0, save=True,
                      gsnap=None):
        check(self.L.lbm_ad_macro_brinkman_collide_natural_convection_p(
            self._prog(), p.h, *[self.f(n) for n in (rho, ux, uy, uz, alpha)], D(nu), q.h,
            *[self.f(n) for n in (tem, qx, qy, qz, diff)], D(gx), D(gy), D(gz), D(tem0), int(save), self.f(gsnap)))

    def ad_collide_bgk(self, p, rho, ux, uy, uz, nu, q, tem, qx, qy, qz, diff, natural, gx=0.0, gy=0.0, gz=0.0,
                       tem0=0.0, save=True):
        a = [self._prog(), p.h, *[self.f(n) for n in (rho, ux, uy, uz)], D(nu), q.h,
             *[self.f(n) for n in (tem, qx, qy, qz)], D(diff)]
        if natural:
            check(self.L.lbm_ad_macro_collide_natural_convection_p(*a, D(gx), D(gy), D(gz), D(tem0), int(save)))
        else:
            check(self.L.lbm_ad_macro_collide_force_convection_p(*a, int(save)))

    def aad_collide_nc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, gx, gy, gz, save=True, igsnap=None):
        check(self.L.lbm_aad_macro_brinkman_collide_natural_convection_p(
            self._prog(), p.h, *[self.f(n) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha)],
            D(nu), q.h, *[self.f(n) for n in (tem, item, iqx, iqy, iqz, diff)], D(gx), D(gy), D(gz), int(save),
            self.f(igsnap)))

    def ans_collide(self, h, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, nu, alpha, save=True):
        check(self.L.lbm_ans_macro_brinkman_collide_p(
            self._prog(), h.h, *[self.f(n) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz)], D(nu),
            self.f(alpha), int(save)))

    def aad_collide_fc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, save=True, igsnap=None):
        check(self.L.lbm_aad_macro_brinkman_collide_force_convection_p(
            self._prog(), p.h, *[self.f(n) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha)],
            D(nu), q.h, *[self.f(n) for n in (tem, item, iqx, iqy, iqz, diff)], int(save), self.f(igsnap)))

    def sens_temp_heat_source(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff,
                              dkds, mask, qn):
        check(self.L.lbm_sensitivity_temperature_at_heat_source(
            q.h, *[self.f(n) for n in (dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff,
                                       dkds)], hptr(self._fi(q, mask)), hptr(self._face(q, qn))))

    def sens_brinkman_diffusivity(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig,
                                  diff, dkds):
        check(self.L.lbm_sensitivity_brinkman_diffusivity(
            q.h, *[self.f(n) for n in (dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff,
                                       dkds)]))

    def residual(self, a, b):
        n = len(a)
        aa = [self.f(x) for x in a] + [None] * (3 - n)
        bb = [self.f(x) for x in b] + [None] * (3 - n)
        N = self.fields[a[0]].size
        out = ctypes.c_double()
        check(self.L.lbm_residual(self.ctx.h, n, *aa, *bb, N, ctypes.byref(out)))
        return out.value


class FusedGpuBackend(GpuBackend):
    """Records stream + boundaries + SmoothCorner and fuses them into the next collide."""

    name = "gpu-fused"

    def __init__(self, ctx=None, prec=pa.F64):
        super().__init__(ctx, prec)
        self._recording = False
        self._pending = None
        self.fused_calls = 0

    def _begin(self):
        if not self._recording:
            check(self.L.lbm_record_begin(self.ctx.h))
            self._recording = True

    def stream(self, h):
        self._begin()
        super().stream(h)

    def istream(self, h):
        self._begin()
        super().istream(h)

    def _prog(self):
        if not self._recording:
            return None
        prog = VP()
        check(self.L.lbm_record_end(self.ctx.h, ctypes.byref(prog)))
        self._recording = False
        if self._pending is not None:
            self.L.lbm_program_destroy(self._pending)
        self._pending = prog
        if self.L.lbm_program_fusable(prog) == 1:
            self.fused_calls += 1
        return prog

    def _flush(self):
        """Run a recorded program that no collide consumed (end of a time loop)."""
        if self._recording:
            prog = VP()
            check(self.L.lbm_record_end(self.ctx.h, ctypes.byref(prog)))
            self._recording = False
            check(self.L.lbm_program_apply(prog))
            self.L.lbm_program_destroy(prog)

    def pops(self, h):
        self._flush()
        return super().pops(h)

    def get(self, name):
        self._flush()
        return super().get(name)

    def residual(self, a, b):
        self._flush()
        return super().residual(a, b)

    def ans_init(self, *a):
        self._flush()
        return super().ans_init(*a)

    def aad_init(self, *a):
        self._flush()
        return super().aad_init(*a)

    def ns_init(self, *a):
        self._flush()
        return super().ns_init(*a)

    def ad_init(self, *a):
        self._flush()
        return super().ad_init(*a)

    def sens_temp_heat_source(self, *a):
        self._flush()
        return super().sens_temp_heat_source(*a)

    def sens_brinkman_diffusivity(self, *a):
        self._flush()
        return super().sens_brinkman_diffusivity(*a)
