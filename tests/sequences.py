"""Driver sequences written once against a small backend interface, so the same time-step loop
runs on the compiled reference harness, on the CPU oracle and on the HIP path (tests/gpu_backend.py).

Fields are addressed by name; each backend owns their storage (numpy here, device memory on the
GPU), which is how the reference drivers pass raw arrays and swap them (heatsink.cpp:170-173).
Boundary lambdas are given as node-indexed arrays.
"""
import ctypes

import numpy as np

from _oracle import D2Q9, D3Q15, NC, ND, Lattice, aos_to_soa, lib, ptr, ref, soa_to_aos, to_face

D = ctypes.c_double


class _FieldStore:
    def __init__(self):
        self.fields = {}

    def put(self, name, arr):
        self.fields[name] = np.array(arr, dtype=np.float64)

    def zeros(self, name, n):
        self.fields[name] = np.zeros(n)

    def get(self, name):
        return self.fields[name].copy()

    def swap(self, a, b):
        self.fields[a], self.fields[b] = self.fields[b], self.fields[a]

    def f(self, name):
        return None if name is None else self.fields[name]


class OracleBackend(_FieldStore):
    name = "oracle"

    def lattice(self, kind, lx, ly, lz=1):
        return Lattice(kind, lx, ly, lz)

    def pops(self, h):
        return h.get()

    def set_pops(self, h, soa):
        h.set(soa)

    def _face(self, h, a, dt=None):
        return to_face(h.kind, h.lx, h.ly, h.lz, a, dt)

    def stream(self, h):
        lib().ol_stream(h.h)

    def istream(self, h):
        lib().ol_istream(h.h)

    def bc(self, h, type_node, adjoint=False):
        lib().ol_bc(h.h, ptr(self._face(h, type_node, np.int32)), int(adjoint))

    def smooth(self, h):
        lib().ol_smooth_corner(h.h)

    def ns_set_u(self, h, mask, ux, uy, uz=None):
        uz = np.zeros_like(ux) if uz is None else uz
        lib().ol_ns_set_u(h.h, ptr(self._face(h, mask, np.int32)), ptr(self._face(h, ux)), ptr(self._face(h, uy)),
                          ptr(self._face(h, uz)))

    def ns_set_rho(self, h, mask, rho, us, ut=None):
        ut = np.zeros_like(us) if ut is None else ut
        lib().ol_ns_set_rho(h.h, ptr(self._face(h, mask, np.int32)), ptr(self._face(h, rho)), ptr(self._face(h, us)),
                            ptr(self._face(h, ut)))

    def ad_set_t(self, h, mask, tem, ux, uy, uz=None):
        lib().ol_ad_set_t(h.h, ptr(self._face(h, mask, np.int32)), ptr(self._face(h, tem)), ptr(self.f(ux)),
                          ptr(self.f(uy)), ptr(self.f(uz)))

    def ad_set_q(self, h, mask, qn, ux, uy, uz, diff):
        field = isinstance(diff, str)
        lib().ol_ad_set_q(h.h, ptr(self._face(h, mask, np.int32)), ptr(self._face(h, qn)), ptr(self.f(ux)),
                          ptr(self.f(uy)), ptr(self.f(uz)), ptr(self.f(diff)) if field else None,
                          D(0.0 if field else float(diff)))

    def ans_iset_u(self, h, mask, ux, uy, uz=None, eps=0.0):
        uz = np.zeros_like(ux) if uz is None else uz
        lib().ol_ans_iset_u(h.h, ptr(self._face(h, mask, np.int32)), ptr(self._face(h, ux)), ptr(self._face(h, uy)),
                            ptr(self._face(h, uz)), D(eps), 1)

    def ans_iset_rho(self, h, mask):
        lib().ol_ans_iset_rho(h.h, ptr(self._face(h, mask, np.int32)))

    def aad_iset_t(self, h, mask, ux, uy, uz=None):
        lib().ol_aad_iset_t(h.h, ptr(self._face(h, mask, np.int32)), ptr(self.f(ux)), ptr(self.f(uy)), ptr(self.f(uz)))

    def aad_iset_q(self, h, mask, ux, uy, uz=None, eps=0.0):
        lib().ol_aad_iset_q(h.h, ptr(self._face(h, mask, np.int32)), ptr(self.f(ux)), ptr(self.f(uy)),
                            ptr(self.f(uz)), D(eps))

    def aad_iset_rho(self, p, q, type_node, rho, ux, uy, uz, tem, eps=0.0):
        lib().ol_aad_iset_rho(p.h, q.h, ptr(self._face(p, type_node, np.int32)), ptr(self.f(rho)), ptr(self.f(ux)),
                              ptr(self.f(uy)), ptr(self.f(uz)), ptr(self.f(tem)), D(eps))

    def ns_init(self, h, rho, ux, uy, uz=None):
        lib().ol_ns_init(h.h, ptr(self.f(rho)), ptr(self.f(ux)), ptr(self.f(uy)), ptr(self.f(uz)))

    def ad_init(self, h, tem, ux, uy, uz=None):
        lib().ol_ad_init(h.h, ptr(self.f(tem)), ptr(self.f(ux)), ptr(self.f(uy)), ptr(self.f(uz)))

    def ans_init(self, h, ux, uy, uz, ip, iux, iuy, iuz):
        lib().ol_ans_init(h.h, *[ptr(self.f(n)) for n in (ux, uy, uz, ip, iux, iuy, iuz)])

    def aad_init(self, h, ux, uy, uz, item, iqx, iqy, iqz):
        lib().ol_aad_init(h.h, *[ptr(self.f(n)) for n in (ux, uy, uz, item, iqx, iqy, iqz)])

    def ns_collide(self, h, rho, ux, uy, uz, nu, save=True):
        lib().ol_ns_macro_collide(h.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz)], D(nu), int(save))

    def ns_brinkman_collide(self, h, rho, ux, uy, uz, nu, alpha, save=True):
        lib().ol_ns_macro_brinkman_collide(h.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz)], D(nu),
                                           ptr(self.f(alpha)), int(save))

    def ad_collide_fc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, save=True, gsnap=None,
                      heat=None):
        lib().ol_ad_collide_fc(p.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz, alpha)], D(nu), q.h,
                               *[ptr(self.f(n)) for n in (tem, qx, qy, qz, diff)], int(save), ptr(self.f(gsnap)),
                               ptr(self.f(heat)))

    def ad_collide_nc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, gx, gy, gz, tem0, save=True,
                      gsnap=None):
        lib().ol_ad_collide_nc(p.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz, alpha)], D(nu), q.h,
                               *[ptr(self.f(n)) for n in (tem, qx, qy, qz, diff)], D(gx), D(gy), D(gz), D(tem0),
                               int(save), ptr(self.f(gsnap)))

    def ad_collide_bgk(self, p, rho, ux, uy, uz, nu, q, tem, qx, qy, qz, diff, natural, gx=0.0, gy=0.0, gz=0.0,
                       tem0=0.0, save=True):
        lib().ol_ad_collide_bgk(p.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz)], D(nu), q.h,
                                *[ptr(self.f(n)) for n in (tem, qx, qy, qz)], D(diff), int(natural), D(gx), D(gy),
                                D(gz), D(tem0), int(save))

    def aad_collide_nc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, gx, gy, gz, save=True, igsnap=None):
        lib().ol_aad_collide_nc(p.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz,
                                                                alpha)], D(nu), q.h,
                                *[ptr(self.f(n)) for n in (tem, item, iqx, iqy, iqz, diff)], D(gx), D(gy), D(gz),
                                int(save), ptr(self.f(igsnap)))

    def ans_collide(self, h, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, nu, alpha, save=True):
        lib().ol_ans_collide(h.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz)],
                             D(nu), ptr(self.f(alpha)), int(save))

    def aad_collide_fc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, save=True, igsnap=None):
        lib().ol_aad_collide_fc(p.h, *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz,
                                                                alpha)], D(nu), q.h,
                                *[ptr(self.f(n)) for n in (tem, item, iqx, iqy, iqz, diff)], int(save),
                                ptr(self.f(igsnap)))

    def sens_temp_heat_source(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff,
                              dkds, mask, qn):
        lib().ol_sens_temp_heat_source(q.h, *[ptr(self.f(n)) for n in (dfds, ux, uy, uz, imx, imy, imz, dads, tem,
                                                                       item, iqx, iqy, iqz, g, ig, diff, dkds)],
                                       ptr(self._face(q, mask, np.int32)), ptr(self._face(q, qn)))

    def sens_brinkman_diffusivity(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig,
                                  diff, dkds):
        lib().ol_sens_brinkman_diffusivity(q.h, *[ptr(self.f(n)) for n in (dfds, ux, uy, uz, imx, imy, imz, dads, tem,
                                                                           item, iqx, iqy, iqz, g, ig, diff, dkds)])

    def residual(self, a, b):
        n = len(a)
        aa = [self.f(x) for x in a] + [None] * (3 - n)
        bb = [self.f(x) for x in b] + [None] * (3 - n)
        N = aa[0].shape[0]
        return lib().ol_residual(n, *[ptr(x) for x in aa], *[ptr(x) for x in bb], N)


class _RefLat:
    def __init__(self, kind, lx, ly, lz):
        self.kind, self.lx, self.ly, self.lz = kind, lx, ly, (lz if ND[kind] == 3 else 1)
        self.N = lx * ly * self.lz
        self.f0 = np.zeros(self.N)
        self.f = np.zeros((NC[kind] - 1) * self.N)

    @property
    def L(self):
        return (self.lx, self.ly, self.lz)

    @property
    def pre(self):
        return "ref2_" if ND[self.kind] == 2 else "ref3_"


class RefBackend(_FieldStore):
    """Each call runs ONE reference function on a freshly built reference lattice (d2q9.h ctor)."""

    name = "reference"

    def _r(self, h, n):
        return getattr(ref(), h.pre + n)

    def lattice(self, kind, lx, ly, lz=1):
        return _RefLat(kind, lx, ly, lz)

    def pops(self, h):
        return aos_to_soa(h.f0, h.f, NC[h.kind])

    def set_pops(self, h, soa):
        h.f0[:], h.f[:] = soa_to_aos(soa)

    def _i(self, a):
        return np.ascontiguousarray(a, dtype=np.int32)

    def _d(self, a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def stream(self, h):
        self._r(h, "stream")(*h.L, ptr(h.f0), ptr(h.f), 0)

    def istream(self, h):
        self._r(h, "stream")(*h.L, ptr(h.f0), ptr(h.f), 1)

    def bc(self, h, type_node, adjoint=False):
        self._r(h, "bc")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(type_node)), int(adjoint))

    def smooth(self, h):
        self._r(h, "smooth")(*h.L, ptr(h.f0), ptr(h.f))

    def ns_set_u(self, h, mask, ux, uy, uz=None):
        uz = np.zeros_like(ux) if uz is None else uz
        self._r(h, "ns_set_u")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self._d(ux)), ptr(self._d(uy)),
                               ptr(self._d(uz)))

    def ns_set_rho(self, h, mask, rho, us, ut=None):
        ut = np.zeros_like(us) if ut is None else ut
        self._r(h, "ns_set_rho")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self._d(rho)), ptr(self._d(us)),
                                 ptr(self._d(ut)))

    def ad_set_t(self, h, mask, tem, ux, uy, uz=None):
        self._r(h, "ad_set_t")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self._d(tem)), ptr(self.f(ux)),
                               ptr(self.f(uy)), ptr(self.f(uz)))

    def ad_set_q(self, h, mask, qn, ux, uy, uz, diff):
        field = isinstance(diff, str)
        self._r(h, "ad_set_q")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self._d(qn)), ptr(self.f(ux)),
                               ptr(self.f(uy)), ptr(self.f(uz)), ptr(self.f(diff)) if field else None,
                               D(0.0 if field else float(diff)))

    def ans_iset_u(self, h, mask, ux, uy, uz=None, eps=0.0):
        uz = np.zeros_like(ux) if uz is None else uz
        self._r(h, "ans_iset_u")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self._d(ux)), ptr(self._d(uy)),
                                 ptr(self._d(uz)), D(eps))

    def ans_iset_rho(self, h, mask):
        self._r(h, "ans_iset_rho")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)))

    def aad_iset_t(self, h, mask, ux, uy, uz=None):
        self._r(h, "aad_iset_t")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self.f(ux)), ptr(self.f(uy)),
                                 ptr(self.f(uz)))

    def aad_iset_q(self, h, mask, ux, uy, uz=None, eps=0.0):
        self._r(h, "aad_iset_q")(*h.L, ptr(h.f0), ptr(h.f), ptr(self._i(mask)), ptr(self.f(ux)), ptr(self.f(uy)),
                                 ptr(self.f(uz)), D(eps))

    def aad_iset_rho(self, p, q, type_node, rho, ux, uy, uz, tem, eps=0.0):
        self._r(p, "aad_iset_rho")(*p.L, ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f), ptr(self._i(type_node)),
                                   ptr(self.f(rho)), ptr(self.f(ux)), ptr(self.f(uy)), ptr(self.f(uz)),
                                   ptr(self.f(tem)), D(eps))

    def ns_init(self, h, rho, ux, uy, uz=None):
        self._r(h, "ns_init")(*h.L, ptr(h.f0), ptr(h.f), *[ptr(self.f(n)) for n in (rho, ux, uy, uz)])

    def ad_init(self, h, tem, ux, uy, uz=None):
        self._r(h, "ad_init")(*h.L, ptr(h.f0), ptr(h.f), *[ptr(self.f(n)) for n in (tem, ux, uy, uz)])

    def ans_init(self, h, ux, uy, uz, ip, iux, iuy, iuz):
        self._r(h, "ans_init")(*h.L, ptr(h.f0), ptr(h.f), *[ptr(self.f(n)) for n in (ux, uy, uz, ip, iux, iuy, iuz)])

    def aad_init(self, h, ux, uy, uz, item, iqx, iqy, iqz):
        self._r(h, "aad_init")(*h.L, ptr(h.f0), ptr(h.f),
                               *[ptr(self.f(n)) for n in (ux, uy, uz, item, iqx, iqy, iqz)])

    def ns_collide(self, h, rho, ux, uy, uz, nu, save=True):
        self._r(h, "ns_macro_collide")(*h.L, ptr(h.f0), ptr(h.f), *[ptr(self.f(n)) for n in (rho, ux, uy, uz)], D(nu),
                                       int(save))

    def ns_brinkman_collide(self, h, rho, ux, uy, uz, nu, alpha, save=True):
        self._r(h, "ns_macro_brinkman_collide")(*h.L, ptr(h.f0), ptr(h.f), *[ptr(self.f(n)) for n in (rho, ux, uy, uz)],
                                                D(nu), ptr(self.f(alpha)), int(save))

    def _snap_out(self, name, q, buf):
        if name is not None:
            self.fields[name] = buf.reshape(q.N, NC[q.kind]).T.copy()

    def ad_collide_fc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, save=True, gsnap=None,
                      heat=None):
        assert heat is None, "the reference has no volumetric heat source"
        buf = np.zeros(NC[q.kind] * q.N)
        self._r(p, "ad_collide_fc")(*p.L, ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f),
                                    *[ptr(self.f(n)) for n in (rho, ux, uy, uz, alpha)], D(nu),
                                    *[ptr(self.f(n)) for n in (tem, qx, qy, qz, diff)], int(save),
                                    ptr(buf) if gsnap else None)
        if save:
            self._snap_out(gsnap, q, buf)

    def ad_collide_bgk(self, p, rho, ux, uy, uz, nu, q, tem, qx, qy, qz, diff, natural, gx=0.0, gy=0.0, gz=0.0,
                       tem0=0.0, save=True):
        args = [ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f), *[ptr(self.f(n)) for n in (rho, ux, uy, uz)], D(nu),
                *[ptr(self.f(n)) for n in (tem, qx, qy, qz)], D(diff)]
        if natural:
            self._r(p, "ad_collide_bgk_nc")(*p.L, *args, D(gx), D(gy), D(gz), D(tem0), int(save))
        else:
            self._r(p, "ad_collide_bgk_fc")(*p.L, *args, int(save))

    def ad_collide_nc(self, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz, diff, gx, gy, gz, tem0, save=True,
                      gsnap=None):
        buf = np.zeros(NC[q.kind] * q.N)
        self._r(p, "ad_collide_nc")(*p.L, ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f),
                                    *[ptr(self.f(n)) for n in (rho, ux, uy, uz, alpha)], D(nu),
                                    *[ptr(self.f(n)) for n in (tem, qx, qy, qz, diff)], D(gx), D(gy), D(gz), D(tem0),
                                    int(save), ptr(buf) if gsnap else None)
        if save:
            self._snap_out(gsnap, q, buf)

    def aad_collide_nc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, gx, gy, gz, save=True, igsnap=None):
        buf = np.zeros(NC[q.kind] * q.N)
        self._r(p, "aad_collide_nc")(*p.L, ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f),
                                     *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz,
                                                                alpha)], D(nu),
                                     *[ptr(self.f(n)) for n in (tem, item, iqx, iqy, iqz, diff)], D(gx), D(gy), D(gz),
                                     int(save), ptr(buf) if igsnap else None)
        if save:
            self._snap_out(igsnap, q, buf)

    def ans_collide(self, h, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, nu, alpha, save=True):
        self._r(h, "ans_collide")(*h.L, ptr(h.f0), ptr(h.f),
                                  *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz)],
                                  D(nu), ptr(self.f(alpha)), int(save))

    def aad_collide_fc(self, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy,
                       iqz, diff, save=True, igsnap=None):
        buf = np.zeros(NC[q.kind] * q.N)
        self._r(p, "aad_collide_fc")(*p.L, ptr(p.f0), ptr(p.f), ptr(q.f0), ptr(q.f),
                                     *[ptr(self.f(n)) for n in (rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz,
                                                                alpha)], D(nu),
                                     *[ptr(self.f(n)) for n in (tem, item, iqx, iqy, iqz, diff)], int(save),
                                     ptr(buf) if igsnap else None)
        if save:
            self._snap_out(igsnap, q, buf)

    def sens_temp_heat_source(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff,
                              dkds, mask, qn):
        assert ND[q.kind] == 2
        gs = np.ascontiguousarray(self.f(g).T).reshape(-1)
        igs = np.ascontiguousarray(self.f(ig).T).reshape(-1)
        ref().ref2_sens_temp_heat_source(q.lx, q.ly, *[ptr(self.f(n)) for n in (dfds, ux, uy, imx, imy, dads, tem,
                                                                               item, iqx, iqy)],
                                         ptr(gs), ptr(igs), ptr(self.f(diff)), ptr(self.f(dkds)),
                                         ptr(self._i(mask)), ptr(self._d(qn)))

    def sens_brinkman_diffusivity(self, q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig,
                                  diff, dkds):
        gs = np.ascontiguousarray(self.f(g).T).reshape(-1)  # reference snapshot layout nc*idx + c
        igs = np.ascontiguousarray(self.f(ig).T).reshape(-1)
        self._r(q, "sens_brinkman_diffusivity")(*q.L, *[ptr(self.f(n)) for n in (dfds, ux, uy, uz, imx, imy, imz,
                                                                                  dads, tem, item, iqx, iqy, iqz)],
                                                ptr(gs), ptr(igs), ptr(self.f(diff)), ptr(self.f(dkds)))

    def residual(self, a, b):
        n = len(a)
        aa = [self.f(x) for x in a] + [None] * (3 - n)
        bb = [self.f(x) for x in b] + [None] * (3 - n)
        fr = ref().ref2_residual
        fr.restype = D
        return fr(n, *[ptr(x) for x in aa], *[ptr(x) for x in bb], aa[0].shape[0])


# ------------------------------------------------------------------------------------------------
class Cavity:
    """test/cavityflow.cpp:44-67 (D2Q9 BGK, bounce-back walls, Zou-He lid, SmoothCorner)."""

    def __init__(self, name, lx, ly, nt, nu=0.1, u0=0.1):
        self.name, self.lx, self.ly, self.nt, self.nu, self.u0 = name, lx, ly, nt, nu, u0

    def masks(self):
        lx, ly = self.lx, self.ly
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        wall = ((i == 0) | (i == lx - 1) | (j == 0)).astype(np.int32)
        lid = (j == ly - 1).astype(np.int32)
        return wall, lid

    def run(self, b):
        N = self.lx * self.ly
        pf = b.lattice(D2Q9, self.lx, self.ly)
        b.put("rho", np.ones(N))
        b.zeros("ux", N)
        b.zeros("uy", N)
        wall, lid = self.masks()
        b.ns_init(pf, "rho", "ux", "uy")
        for _ in range(self.nt):
            b.ns_collide(pf, "rho", "ux", "uy", None, self.nu, True)
            b.stream(pf)
            b.bc(pf, wall)
            b.ns_set_u(pf, lid, np.full(N, self.u0), np.zeros(N))
            b.smooth(pf)
        return {"f": b.pops(pf), "rho": b.get("rho"), "ux": b.get("ux"), "uy": b.get("uy")}


class ForcedConvection2D:
    """Coupled NS + AD forced convection in a channel, forward then adjoint, then the heat-source
    sensitivity: the call order of test/nsadsens.cpp:67-133 with the ForceConvection collides
    (advection.h:814, adjointadvection.h:997) and the heat-flux objective of
    production/heatsink.cpp:156-166,200-232."""

    def __init__(self, name, lx, ly, nt, nu=0.1, pr=6.0, umax=0.05, qn=1e-2, L=None, seed=5, check=None, swap=True):
        self.name, self.lx, self.ly, self.nt = name, lx, ly, nt
        self.nu, self.pr, self.umax, self.qn = nu, pr, umax, qn
        self.L = L if L is not None else lx // 2
        self.seed = seed
        self.check = check
        # swap = False: one array per field, as test/nsadsens.cpp:70-101 (the boundary calls then
        # read the very arrays the next collide saves into)
        self.swap_fields = swap

    def setup(self):
        lx, ly = self.lx, self.ly
        N = lx * ly
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        rng = np.random.default_rng(self.seed)
        s = rng.uniform(0.0, 1.0, N)
        qf, qg, amax = 0.01, 1.0, 1.0
        dfl = self.nu / self.pr
        dso = 10.0 * dfl
        d = dict(
            i=i, j=j, s=s,
            alpha=amax * qf * (1.0 - s) / (s + qf),
            dads=-amax * qf * (1.0 + qf) / (s + qf) ** 2,
            diff=dso + (dfl - dso) * s * (1.0 + qg) / (s + qg),
            dkds=(dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2,
        )
        H = ly - 1.0
        d["uin"] = 4.0 * self.umax * j * (H - j) / H ** 2
        d["wall_f"] = np.where(j == 0, 1, np.where(j == ly - 1, 2, 0)).astype(np.int32)
        d["inlet"] = ((i == 0) & (j > 0) & (j < ly - 1)).astype(np.int32)
        d["outlet"] = ((i == lx - 1) & (j > 0) & (j < ly - 1)).astype(np.int32)
        d["wall_g"] = np.where(j == ly - 1, 2, 0).astype(np.int32)
        d["heat_edge"] = ((j == 0) & (i < self.L)).astype(np.int32)
        d["flux_bc"] = (((j == 0) | (i == lx - 1)) & ~((i == 0) & (j > 0))).astype(np.int32)
        d["qn"] = np.where((j == 0) & (i < self.L), self.qn, 0.0)
        d["outlet_type"] = np.where(d["outlet"] == 1, 2, 0).astype(np.int32)
        return d

    def run(self, b):
        lx, ly, N = self.lx, self.ly, self.lx * self.ly
        d = self.setup()
        pf = b.lattice(D2Q9, lx, ly)
        pg = b.lattice(D2Q9, lx, ly)
        for k in ("alpha", "dads", "diff", "dkds"):
            b.put(k, d[k])
        b.put("rho", np.ones(N))
        for k in ("ux", "uy", "uxp", "uyp", "tem", "qx", "qy", "qxp", "qyp", "ip", "iux", "iuy", "iuxp", "iuyp",
                  "imx", "imy", "item", "iqx", "iqy", "iqxp", "iqyp", "dfds"):
            b.zeros(k, N)
        b.zeros("g", (NC[D2Q9], N))
        b.zeros("ig", (NC[D2Q9], N))
        zero = np.zeros(N)
        b.ns_init(pf, "rho", "ux", "uy")
        b.ad_init(pg, "tem", "ux", "uy")
        res = []
        for t in range(1, self.nt + 1):
            b.ad_collide_fc(pf, "rho", "ux", "uy", None, "alpha", self.nu, pg, "tem", "qx", "qy", None, "diff",
                            True, "g")
            if self.check and t % self.check == 0:
                res.append((b.residual(["ux", "uy"], ["uxp", "uyp"]), b.residual(["qx", "qy"], ["qxp", "qyp"])))
            b.stream(pf)
            b.stream(pg)
            b.bc(pf, d["wall_f"])
            b.ns_set_u(pf, d["inlet"], d["uin"], zero)
            b.ns_set_rho(pf, d["outlet"], np.ones(N), zero)
            b.smooth(pf)
            b.bc(pg, d["wall_g"])
            b.ad_set_t(pg, d["inlet"], zero, "ux", "uy")
            b.ad_set_q(pg, d["flux_bc"], d["qn"], "ux", "uy", None, "diff")
            b.smooth(pg)
            if self.swap_fields:
                for v in ("ux", "uy", "qx", "qy"):
                    b.swap(v, v + "p")
        # after the final swap the step-nt fields sit in the "p" buffers (heatsink.cpp:170-173)
        if self.swap_fields:
            for v in ("ux", "uy", "qx", "qy"):
                b.swap(v, v + "p")
        out = {"f": b.pops(pf), "g": b.pops(pg)}
        for k in ("rho", "ux", "uy", "tem", "qx", "qy", "g"):
            out["fwd_" + k] = b.get(k)
        b.ans_init(pf, "ux", "uy", None, "ip", "iux", "iuy", None)
        b.aad_init(pg, "ux", "uy", None, "item", "iqx", "iqy", None)
        for t in range(1, self.nt + 1):
            b.aad_collide_fc(pf, "rho", "ux", "uy", None, "ip", "iux", "iuy", None, "imx", "imy", None, "alpha",
                             self.nu, pg, "tem", "item", "iqx", "iqy", None, "diff", True, "ig")
            if self.check and t % self.check == 0:
                res.append((b.residual(["iux", "iuy"], ["iuxp", "iuyp"]), b.residual(["iqx", "iqy"], ["iqxp", "iqyp"])))
            b.istream(pf)
            b.istream(pg)
            b.aad_iset_t(pg, d["inlet"], "ux", "uy")
            b.aad_iset_q(pg, d["flux_bc"], "ux", "uy")
            b.aad_iset_q(pg, d["heat_edge"], "ux", "uy", None, 1.0)
            b.bc(pg, d["wall_g"], adjoint=True)
            b.smooth(pg)
            b.bc(pf, d["wall_f"], adjoint=True)
            b.ans_iset_u(pf, d["inlet"], d["uin"], zero)
            b.aad_iset_rho(pf, pg, d["outlet_type"], "rho", "ux", "uy", None, "tem")
            b.smooth(pf)
            if self.swap_fields:
                for v in ("iux", "iuy", "iqx", "iqy"):
                    b.swap(v, v + "p")
        if self.swap_fields:
            for v in ("iux", "iuy", "iqx", "iqy"):
                b.swap(v, v + "p")
        b.sens_temp_heat_source(pg, "dfds", "ux", "uy", None, "imx", "imy", None, "dads", "tem", "item", "iqx", "iqy",
                                None, "g", "ig", "diff", "dkds", d["heat_edge"], d["qn"])
        out.update({"adj_f": b.pops(pf), "adj_g": b.pops(pg), "dfds": b.get("dfds")})
        for k in ("ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy", "ig"):
            out["adj_" + k] = b.get(k)
        tem = out["fwd_tem"]
        out["objective"] = np.array(sum(tem[i] for i in range(lx) if i < self.L) / self.L)
        if res:
            out["residuals"] = np.array(res)
        return out


class ForcedConvection3D:
    """D3Q15 NS + D3Q15 AD forced convection in a duct, forward then adjoint, then the Brinkman and
    diffusivity sensitivity: the call order of test/nsadsens.cpp:67-133 on the 3D overloads
    (advection.h:871, adjointadvection.h:1054, adjointadvection.h:1478).  Walls y = 0, ly-1 and
    z = 0, lz-1 bounce back; inlet x = 0 Zou-He velocity, outlet x = lx-1 Zou-He pressure; T = 0 at
    the inlet, heat flux into the y = 0 wall for x < L, adiabatic elsewhere.  The adjoint outlet is
    ANS::iBoundaryConditionSetRho3D (the reference's 3D AAD::iBoundaryConditionSetRho does not
    compile)."""

    def __init__(self, name, lx, ly, lz, nt, nu=0.1, pr=6.0, umax=0.05, qn=1e-2, seed=6, check=None,
                 kinds=(D3Q15, D3Q15)):
        self.name, self.lx, self.ly, self.lz, self.nt = name, lx, ly, lz, nt
        self.nu, self.pr, self.umax, self.qn = nu, pr, umax, qn
        self.L = lx // 2
        self.seed = seed
        self.check = check
        self.kinds = kinds  # (NS lattice, AD lattice); the reference's is (D3Q15, D3Q15)

    def setup(self):
        lx, ly, lz = self.lx, self.ly, self.lz
        N = lx * ly * lz
        n = np.arange(N)
        i, j, k = n % lx, (n // lx) % ly, n // (lx * ly)
        rng = np.random.default_rng(self.seed)
        s = rng.uniform(0.0, 1.0, N)
        qf, qg, amax = 0.01, 1.0, 1.0
        dfl = self.nu / self.pr
        dso = 10.0 * dfl
        d = dict(
            alpha=amax * qf * (1.0 - s) / (s + qf),
            dads=-amax * qf * (1.0 + qf) / (s + qf) ** 2,
            diff=dso + (dfl - dso) * s * (1.0 + qg) / (s + qg),
            dkds=(dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2,
        )
        Hy, Hz = ly - 1.0, lz - 1.0
        d["uin"] = 16.0 * self.umax * j * (Hy - j) * k * (Hz - k) / (Hy ** 2 * Hz ** 2)
        side = (j == 0) | (j == ly - 1) | (k == 0) | (k == lz - 1)
        inner = (j > 0) & (j < ly - 1) & (k > 0) & (k < lz - 1)
        d["wall_f"] = side.astype(np.int32)
        d["inlet"] = ((i == 0) & inner).astype(np.int32)
        d["outlet"] = ((i == lx - 1) & inner).astype(np.int32)
        d["wall_g"] = np.where(j == ly - 1, 2, 0).astype(np.int32)
        d["heat_face"] = ((j == 0) & (i < self.L)).astype(np.int32)
        d["flux_bc"] = (((j == 0) | (k == 0) | (k == lz - 1) | (i == lx - 1)) & (i > 0)).astype(np.int32)
        d["qn"] = np.where((j == 0) & (i < self.L), self.qn, 0.0)
        return d

    def run(self, b):
        lx, ly, lz = self.lx, self.ly, self.lz
        N = lx * ly * lz
        d = self.setup()
        pf = b.lattice(self.kinds[0], lx, ly, lz)
        pg = b.lattice(self.kinds[1], lx, ly, lz)
        for k in ("alpha", "dads", "diff", "dkds"):
            b.put(k, d[k])
        b.put("rho", np.ones(N))
        for k in ("ux", "uy", "uz", "uxp", "uyp", "uzp", "tem", "qx", "qy", "qz", "qxp", "qyp", "qzp",
                  "ip", "iux", "iuy", "iuz", "iuxp", "iuyp", "iuzp", "imx", "imy", "imz",
                  "item", "iqx", "iqy", "iqz", "iqxp", "iqyp", "iqzp", "dfds"):
            b.zeros(k, N)
        b.zeros("g", (NC[self.kinds[1]], N))
        b.zeros("ig", (NC[self.kinds[1]], N))
        zero = np.zeros(N)
        b.ns_init(pf, "rho", "ux", "uy", "uz")
        b.ad_init(pg, "tem", "ux", "uy", "uz")
        res = []
        vel, flx = ("ux", "uy", "uz"), ("qx", "qy", "qz")
        for t in range(1, self.nt + 1):
            b.ad_collide_fc(pf, "rho", "ux", "uy", "uz", "alpha", self.nu, pg, "tem", "qx", "qy", "qz", "diff",
                            True, "g")
            if self.check and t % self.check == 0:
                res.append((b.residual(list(vel), [v + "p" for v in vel]), b.residual(list(flx), [v + "p" for v in flx])))
            b.stream(pf)
            b.stream(pg)
            b.bc(pf, d["wall_f"])
            b.ns_set_u(pf, d["inlet"], d["uin"], zero, zero)
            b.ns_set_rho(pf, d["outlet"], np.ones(N), zero, zero)
            b.smooth(pf)
            b.bc(pg, d["wall_g"])
            b.ad_set_t(pg, d["inlet"], zero, "ux", "uy", "uz")
            b.ad_set_q(pg, d["flux_bc"], d["qn"], "ux", "uy", "uz", "diff")
            b.smooth(pg)
            for v in vel + flx:
                b.swap(v, v + "p")
        for v in vel + flx:
            b.swap(v, v + "p")
        out = {"f": b.pops(pf), "g": b.pops(pg)}
        for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz", "g"):
            out["fwd_" + k] = b.get(k)
        b.ans_init(pf, "ux", "uy", "uz", "ip", "iux", "iuy", "iuz")
        b.aad_init(pg, "ux", "uy", "uz", "item", "iqx", "iqy", "iqz")
        ivel, iflx = ("iux", "iuy", "iuz"), ("iqx", "iqy", "iqz")
        for t in range(1, self.nt + 1):
            b.aad_collide_fc(pf, "rho", "ux", "uy", "uz", "ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "alpha",
                             self.nu, pg, "tem", "item", "iqx", "iqy", "iqz", "diff", True, "ig")
            if self.check and t % self.check == 0:
                res.append((b.residual(list(ivel), [v + "p" for v in ivel]),
                            b.residual(list(iflx), [v + "p" for v in iflx])))
            b.istream(pf)
            b.istream(pg)
            b.aad_iset_t(pg, d["inlet"], "ux", "uy", "uz")
            b.aad_iset_q(pg, d["flux_bc"], "ux", "uy", "uz")
            b.aad_iset_q(pg, d["heat_face"], "ux", "uy", "uz", 1.0)
            b.bc(pg, d["wall_g"], adjoint=True)
            b.smooth(pg)
            b.bc(pf, d["wall_f"], adjoint=True)
            b.ans_iset_u(pf, d["inlet"], d["uin"], zero, zero)
            b.ans_iset_rho(pf, d["outlet"])
            b.smooth(pf)
            for v in ivel + iflx:
                b.swap(v, v + "p")
        for v in ivel + iflx:
            b.swap(v, v + "p")
        b.sens_brinkman_diffusivity(pg, "dfds", "ux", "uy", "uz", "imx", "imy", "imz", "dads", "tem", "item", "iqx",
                                    "iqy", "iqz", "g", "ig", "diff", "dkds")
        out.update({"adj_f": b.pops(pf), "adj_g": b.pops(pg), "dfds": b.get("dfds")})
        # the 3D adjoint collide leaves iuz as it found it (adjointadvection.h:1074-1084)
        for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz", "ig"):
            out["adj_" + k] = b.get(k)
        if res:
            out["residuals"] = np.array(res)
        return out


class NaturalConvection2D:
    """Coupled NS + AD natural convection, forward then adjoint, then the heat-source sensitivity: the
    call order of test/nsadncsens.cpp:87-177 (MacroBrinkmanCollideNaturalConvection, advection.h:930
    and adjointadvection.h:1113; mirror at x = 0, bounce-back elsewhere; T = tem0 on x = lx-1 and
    y = ly-1; heat flux qn into y = 0 for x < L, adiabatic elsewhere on y = 0).  As in that driver the
    velocity and flux arrays are swapped after every step and not swapped back, so the adjoint reads
    the last-but-one forward velocity.  gx is nonzero here (0 in the driver) so that both buoyancy
    components act; Ra is lowered so that the small cavity stays stable."""

    def __init__(self, name, lx, ly, nt, nu=0.1, pr=6.0, ra=3.0e3, qn=1e-2, L=2.0, seed=7, check=None):
        self.name, self.lx, self.ly, self.nt = name, lx, ly, nt
        self.nu, self.pr, self.ra, self.qn, self.L = nu, pr, ra, qn, L
        self.seed, self.check = seed, check

    def setup(self):
        lx, ly = self.lx, self.ly
        N = lx * ly
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        rng = np.random.default_rng(self.seed)
        s = rng.uniform(0.01, 0.99, N)
        qf, qg, amax = 1e-2, 1e0, 1.0e4
        dfl = self.nu / self.pr
        dso = 10.0 * dfl
        u = self.nu * np.sqrt(self.ra / self.pr) / (ly - 1.0)
        d = dict(
            alpha=amax / (ly - 1.0) * qf * (1.0 - s) / (s + qf),
            dads=-amax / (ly - 1.0) * qf * (1.0 + qf) / (s + qf) ** 2,
            diff=dso + (dfl - dso) * s * (1.0 + qg) / (s + qg),
            dkds=(dfl - dso) * qg * (1.0 + qg) / (s + qg) ** 2,
        )
        self.gy = u * u / (ly - 1.0)
        self.gx = 0.3 * self.gy
        self.tem0 = 0.0
        d["type_f"] = np.where(i == 0, 2, 1).astype(np.int32)
        d["type_g"] = np.where(i == 0, 2, 0).astype(np.int32)
        d["set_t"] = ((i == lx - 1) | (j == ly - 1)).astype(np.int32)
        d["bottom"] = (j == 0).astype(np.int32)
        d["heat_edge"] = ((j == 0) & (i < self.L)).astype(np.int32)
        d["qn"] = np.where((j == 0) & (i < self.L), self.qn, 0.0)
        return d

    def run(self, b):
        lx, ly, N = self.lx, self.ly, self.lx * self.ly
        d = self.setup()
        pf = b.lattice(D2Q9, lx, ly)
        pg = b.lattice(D2Q9, lx, ly)
        for k in ("alpha", "dads", "diff", "dkds"):
            b.put(k, d[k])
        b.put("rho", np.ones(N))
        for k in ("ux", "uy", "uxp", "uyp", "tem", "qx", "qy", "qxp", "qyp", "ip", "iux", "iuy", "iuxp", "iuyp",
                  "imx", "imy", "item", "iqx", "iqy", "iqxp", "iqyp", "dfds"):
            b.zeros(k, N)
        b.zeros("g", (NC[D2Q9], N))
        b.zeros("ig", (NC[D2Q9], N))
        tem0 = np.full(N, self.tem0)
        g = (self.gx, self.gy, 0.0)
        b.ns_init(pf, "rho", "ux", "uy")
        b.ad_init(pg, "tem", "ux", "uy")
        res = []
        for t in range(1, self.nt + 1):
            b.ad_collide_nc(pf, "rho", "ux", "uy", None, "alpha", self.nu, pg, "tem", "qx", "qy", None, "diff", *g,
                            self.tem0, True, "g")
            if self.check and t % self.check == 0:
                res.append((b.residual(["ux", "uy"], ["uxp", "uyp"]), b.residual(["qx", "qy"], ["qxp", "qyp"])))
            b.stream(pf)
            b.stream(pg)
            b.bc(pf, d["type_f"])
            b.ad_set_t(pg, d["set_t"], tem0, "ux", "uy")
            b.ad_set_q(pg, d["bottom"], d["qn"], "ux", "uy", None, "diff")
            b.bc(pg, d["type_g"])
            b.smooth(pf)
            b.smooth(pg)
            for v in ("ux", "uy", "qx", "qy"):
                b.swap(v, v + "p")
        out = {"f": b.pops(pf), "g": b.pops(pg)}
        for k in ("rho", "ux", "uy", "uxp", "uyp", "tem", "qx", "qy", "g"):
            out["fwd_" + k] = b.get(k)
        b.ans_init(pf, "ux", "uy", None, "ip", "iux", "iuy", None)
        b.aad_init(pg, "ux", "uy", None, "item", "iqx", "iqy", None)
        for t in range(1, self.nt + 1):
            b.aad_collide_nc(pf, "rho", "ux", "uy", None, "ip", "iux", "iuy", None, "imx", "imy", None, "alpha",
                             self.nu, pg, "tem", "item", "iqx", "iqy", None, "diff", *g, True, "ig")
            if self.check and t % self.check == 0:
                res.append((b.residual(["iux", "iuy"], ["iuxp", "iuyp"]), b.residual(["iqx", "iqy"], ["iqxp", "iqyp"])))
            b.istream(pf)
            b.istream(pg)
            b.aad_iset_t(pg, d["set_t"], "ux", "uy")
            b.aad_iset_q(pg, d["bottom"], "ux", "uy")
            b.aad_iset_q(pg, d["heat_edge"], "ux", "uy", None, 1.0)
            b.bc(pg, d["type_g"], adjoint=True)
            b.bc(pf, d["type_f"], adjoint=True)
            b.smooth(pf)
            b.smooth(pg)
            for v in ("iux", "iuy", "iqx", "iqy"):
                b.swap(v, v + "p")
        b.sens_temp_heat_source(pg, "dfds", "ux", "uy", None, "imx", "imy", None, "dads", "tem", "item", "iqx", "iqy",
                                None, "g", "ig", "diff", "dkds", d["heat_edge"], d["qn"])
        out.update({"adj_f": b.pops(pf), "adj_g": b.pops(pg), "dfds": b.get("dfds")})
        for k in ("ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy", "ig"):
            out["adj_" + k] = b.get(k)
        tem = out["fwd_tem"]
        out["objective"] = np.array(sum(tem[i] for i in range(lx) if i < self.L) / self.L)
        if res:
            out["residuals"] = np.array(res)
        return out


class TransientNaturalConvection2D(NaturalConvection2D):
    """production/heatsink_transient.cpp:138-222: the transient natural-convection heat sink, its
    objective (the heat-source edge temperature averaged over every step) and the adjoint run
    backward in time over the stored forward trajectory, accumulating the heat-source sensitivity
    at every step.  Geometry, boundaries and coefficients as NaturalConvection2D (the same driver
    family); D2Q9 for both lattices as in that program; T starts at tem0 = 0.  The program leaves the
    step-0 populations snapshot gi[0] unset before the adjoint's last step reads it; here it holds
    the initial populations (AD::InitialCondition's equilibrium)."""

    def run(self, b):
        lx, ly, N, nt = self.lx, self.ly, self.lx * self.ly, self.nt
        d = self.setup()
        pf = b.lattice(D2Q9, lx, ly)
        pg = b.lattice(D2Q9, lx, ly)
        for k in ("alpha", "dads", "diff", "dkds"):
            b.put(k, d[k])
        b.put("rho_0", np.ones(N))
        for k in ("ux_0", "uy_0", "tem_0", "qx", "qy"):
            b.zeros(k, N)
        tem0 = np.full(N, self.tem0)
        g = (self.gx, self.gy, 0.0)
        b.ns_init(pf, "rho_0", "ux_0", "uy_0")
        b.ad_init(pg, "tem_0", "ux_0", "uy_0")
        b.put("gi_0", b.pops(pg))
        for t in range(1, nt):
            for k in ("rho", "ux", "uy", "tem"):
                b.zeros("%s_%d" % (k, t), N)
            b.zeros("gi_%d" % t, (NC[D2Q9], N))
            ux, uy = "ux_%d" % t, "uy_%d" % t
            b.ad_collide_nc(pf, "rho_%d" % t, ux, uy, None, "alpha", self.nu, pg, "tem_%d" % t, "qx", "qy", None, "diff",
                            *g, self.tem0, True, "gi_%d" % t)
            b.stream(pf)
            b.stream(pg)
            b.bc(pf, d["type_f"])
            b.ad_set_t(pg, d["set_t"], tem0, ux, uy)
            b.ad_set_q(pg, d["bottom"], d["qn"], ux, uy, None, "diff")
            b.bc(pg, d["type_g"])
            b.smooth(pf)
            b.smooth(pg)
        out = {"f": b.pops(pf), "g": b.pops(pg)}
        for k in ("ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy", "dfds"):
            b.zeros(k, N)
        b.zeros("ig", (NC[D2Q9], N))
        last = nt - 1
        b.ans_init(pf, "ux_%d" % last, "uy_%d" % last, None, "ip", "iux", "iuy", None)
        b.aad_init(pg, "ux_%d" % last, "uy_%d" % last, None, "item", "iqx", "iqy", None)
        for t in range(nt - 2, -1, -1):
            ux, uy = "ux_%d" % t, "uy_%d" % t
            b.aad_collide_nc(pf, "rho_%d" % t, ux, uy, None, "ip", "iux", "iuy", None, "imx", "imy", None, "alpha",
                             self.nu, pg, "tem_%d" % t, "item", "iqx", "iqy", None, "diff", *g, True, "ig")
            b.sens_temp_heat_source(pg, "dfds", ux, uy, None, "imx", "imy", None, "dads", "tem_%d" % t, "item", "iqx",
                                    "iqy", None, "gi_%d" % t, "ig", "diff", "dkds", d["heat_edge"], d["qn"])
            b.istream(pf)
            b.istream(pg)
            b.aad_iset_t(pg, d["set_t"], ux, uy)
            b.aad_iset_q(pg, d["bottom"], ux, uy)
            b.aad_iset_q(pg, d["heat_edge"], ux, uy, None, 1.0)
            b.bc(pg, d["type_g"], adjoint=True)
            b.bc(pf, d["type_f"], adjoint=True)
            b.smooth(pf)
            b.smooth(pg)
        out.update({"adj_f": b.pops(pf), "adj_g": b.pops(pg), "dfds": b.get("dfds")})
        for k in ("ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy", "ig"):
            out["adj_" + k] = b.get(k)
        edge = [i for i in range(lx) if i < self.L]
        obj = 0.0
        for t in range(nt):
            tem = b.get("tem_%d" % t)
            for i in edge:
                obj += tem[i]
        out["objective"] = np.array(obj / (self.L * nt))
        out["tem_last"] = b.get("tem_%d" % last)
        return out


class NaturalConvectionBGK2D:
    """test/naturalconvection.cpp:30-63: D2Q9 flow and D2Q9 temperature, the BGK-only
    AD::MacroCollideNaturalConvection (advection.h:616-662) saving every step; mirror walls at y = 0
    and y = ly-1 and periodic in x for the flow, periodic for the temperature, T = Th at y = 0 and
    Tl at y = ly-1 (AD::BoundaryConditionSetT), SmoothCorner on both.  The program's 101x51 box and
    1e5 steps shrink to a small box; gravity is raised so that the layer (Ra ~ 2e4) overturns within
    the steps run, and a small seeded perturbation of the initial temperature breaks the symmetry
    (the program starts from T = (Th + Tl)/2 exactly)."""

    def __init__(self, name, lx, ly, nt, nu=0.02, alpha=0.02, gy=1.0e-3, th=2.0, tl=1.0, seed=9):
        self.name, self.lx, self.ly, self.nt = name, lx, ly, nt
        self.nu, self.alpha, self.gy, self.th, self.tl = nu, alpha, gy, th, tl
        self.seed = seed

    def run(self, b):
        lx, ly, N = self.lx, self.ly, self.lx * self.ly
        i, j = np.meshgrid(np.arange(lx), np.arange(ly), indexing="xy")
        i, j = i.reshape(-1), j.reshape(-1)
        wall = (j == 0) | (j == ly - 1)
        type_f = np.where(wall, 2, 0).astype(np.int32)
        type_g = np.zeros(N, dtype=np.int32)
        tem_b = np.where(j == 0, self.th, np.where(j == ly - 1, self.tl, 0.0))
        tm = 0.5 * (self.th + self.tl)
        rng = np.random.default_rng(self.seed)
        pf = b.lattice(D2Q9, lx, ly)
        pg = b.lattice(D2Q9, lx, ly)
        b.put("rho", np.ones(N))
        b.put("tem", tm + 0.01 * rng.uniform(-1, 1, N))
        for k in ("ux", "uy", "qx", "qy"):
            b.zeros(k, N)
        b.ns_init(pf, "rho", "ux", "uy")
        b.ad_init(pg, "tem", "ux", "uy")
        for _ in range(self.nt):
            b.ad_collide_bgk(pf, "rho", "ux", "uy", None, self.nu, pg, "tem", "qx", "qy", None, self.alpha, 1,
                             0.0, self.gy, 0.0, tm, True)
            b.stream(pf)
            b.stream(pg)
            b.bc(pf, type_f)
            b.bc(pg, type_g)
            b.ad_set_t(pg, wall.astype(np.int32), tem_b, "ux", "uy")
            b.smooth(pf)
            b.smooth(pg)
        out = {"f": b.pops(pf), "g": b.pops(pg)}
        for k in ("rho", "ux", "uy", "tem", "qx", "qy"):
            out[k] = b.get(k)
        return out


SEQUENCES = [
    Cavity("cavity_21x21", 21, 21, 300),
    # the lid-driven cavity workload: 101 x 101 nodes (L = 100, Re = u0 L / nu = 100), 10,000 steps
    Cavity("cavity_101x101", 101, 101, 10000),
    ForcedConvection2D("fc2d_24x16", 24, 16, 200, check=100),
    ForcedConvection3D("fc3d_7x5x5", 7, 5, 5, 60, check=30),
    ForcedConvection2D("fc2d_noswap_20x12", 20, 12, 120, swap=False),
    NaturalConvection2D("nc2d_22x18", 22, 18, 200, check=100),
    NaturalConvectionBGK2D("ncbgk2d_32x17", 32, 17, 600),
    TransientNaturalConvection2D("transient_nc2d_16x12", 16, 12, 60, L=3.0),
]
