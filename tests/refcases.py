"""Case definitions shared by the golden-fixture generator (tests/golden/make_golden.py) and the
CPU oracle tests.  Each case has seeded inputs, a runner on the compiled reference harness
(reference AoS layout, node-indexed lambdas) and a runner on the oracle (SoA layout, face
arrays).  Both runners return the same dict of outputs.

Reference functions exercised are named in each case's `cite`.
"""
import ctypes

import numpy as np

from _oracle import (D2Q9, D3Q7, D3Q15, D3Q27, NC, ND, Lattice, aos_to_soa, lib, ptr, ref, ref_avx, soa_to_aos,
                     to_face)

D = ctypes.c_double


def _pops(rng, kind, N, scale=0.05):
    """Populations near a rest equilibrium with rho ~ 1 (all positive)."""
    nc = NC[kind]
    w = {9: [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4,
         15: [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8,
         27: [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8,
         7: [0.05] + [1 / 6] * 6}[nc]  # D3Q7's rest weight is 0; give its population a value anyway
    return np.array(w)[:, None] * (1.0 + scale * rng.uniform(-1, 1, (nc, N)))


def _fields(rng, N):
    return dict(
        rho=1.0 + 0.02 * rng.uniform(-1, 1, N),
        ux=0.05 * rng.uniform(-1, 1, N),
        uy=0.05 * rng.uniform(-1, 1, N),
        uz=0.05 * rng.uniform(-1, 1, N),
        tem=rng.uniform(0, 1, N),
        alpha=rng.uniform(0, 5, N),
        diff=rng.uniform(0.01, 0.2, N),
        dads=rng.uniform(-3, 0, N),
        dkds=rng.uniform(-0.1, 0.1, N),
    )


# buoyancy parameters of the natural-convection cases: every component and T - tem0 nonzero
NC_G = (1.3e-3, -2.1e-3, 1.7e-3)
NC_TEM0 = 0.4


def he_inputs(inp):
    """Heat-exchange coefficient beta (per node), its design derivative and the constant
    diffusivity, derived from the seeded inputs."""
    return 0.1 * inp["alpha"], inp["dkds"], float(inp["dconst"])


def mf_dirs(inp):
    """Direction fields of the mass-flow objective, O(1) and derived from the seeded inputs."""
    return 10.0 * inp["v0"], 10.0 * inp["v1"], 10.0 * inp["v2"]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Case:
    def __init__(self, name, kind, shape, seed, cite, kind_b=None):
        self.name, self.kind, self.seed, self.cite = name, kind, seed, cite
        self.kind_b = kind if kind_b is None else kind_b  # the advection-diffusion lattice
        self.lx, self.ly, self.lz = shape if len(shape) == 3 else (shape[0], shape[1], 1)
        self.N = self.lx * self.ly * self.lz

    def face(self, node_arr, dtype=None):
        return to_face(self.kind, self.lx, self.ly, self.lz, node_arr, dtype)

    @property
    def rk(self):
        """reference harness symbol prefix"""
        return "ref2_" if ND[self.kind] == 2 else "ref3_"

    def dims(self):
        return (self.lx, self.ly, self.lz)


# ----------------------------------------------------------------------------------------------
# population-only operators: stream, lattice BCs, corner smoothing
# ----------------------------------------------------------------------------------------------
def make_inputs(case):
    rng = np.random.default_rng(case.seed)
    inp = {"f": _pops(rng, case.kind, case.N), "g": _pops(rng, case.kind_b, case.N)}
    inp.update(_fields(rng, case.N))
    inp["type"] = rng.integers(0, 3, case.N).astype(np.int32)
    inp["mask"] = (rng.uniform(size=case.N) < 0.7).astype(np.int32)
    inp["v0"] = 0.05 * rng.uniform(-1, 1, case.N)
    inp["v1"] = 0.05 * rng.uniform(-1, 1, case.N)
    inp["v2"] = 0.05 * rng.uniform(-1, 1, case.N)
    inp["rhob"] = 1.0 + 0.02 * rng.uniform(-1, 1, case.N)
    inp["eps"] = np.array(rng.uniform(0.1, 1.0))
    inp["dconst"] = np.array(rng.uniform(0.02, 0.1))
    inp["dfds0"] = rng.uniform(-1, 1, case.N)
    return inp


def _ref_pops(case, soa):
    f0, f = soa_to_aos(soa)
    return np.ascontiguousarray(f0), np.ascontiguousarray(f)


def run_ref(case, op, inp):
    R = ref()
    L = (case.lx, case.ly, case.lz)
    nc = NC[case.kind]
    f0, f = _ref_pops(case, inp["f"])
    g0, g = _ref_pops(case, inp["g"])
    fn = lambda n: getattr(R, case.rk + n)
    out = {}
    uz = inp["uz"] if ND[case.kind] == 3 else None
    if op in ("stream", "istream"):
        fn("stream")(*L, ptr(f0), ptr(f), int(op == "istream"))
    elif op in ("bc", "ibc"):
        fn("bc")(*L, ptr(f0), ptr(f), ptr(inp["type"]), int(op == "ibc"))
    elif op == "smooth":
        fn("smooth")(*L, ptr(f0), ptr(f))
    elif op == "ns_set_u":
        fn("ns_set_u")(*L, ptr(f0), ptr(f), ptr(inp["mask"]), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]))
    elif op == "ns_set_rho":
        fn("ns_set_rho")(*L, ptr(f0), ptr(f), ptr(inp["mask"]), ptr(inp["rhob"]), ptr(inp["v0"]), ptr(inp["v1"]))
    elif op == "ans_iset_u":
        fn("ans_iset_u")(*L, ptr(f0), ptr(f), ptr(inp["mask"]), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]),
                         D(float(inp["eps"])))
    elif op == "ans_iset_rho":
        fn("ans_iset_rho")(*L, ptr(f0), ptr(f), ptr(inp["mask"]))
    elif op == "ad_set_t":
        fn("ad_set_t")(*L, ptr(g0), ptr(g), ptr(inp["mask"]), ptr(inp["tem"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ad_set_q_field":
        fn("ad_set_q")(*L, ptr(g0), ptr(g), ptr(inp["mask"]), ptr(inp["v0"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                       ptr(inp["diff"]), D(0.0))
    elif op == "ad_set_q_const":
        fn("ad_set_q")(*L, ptr(g0), ptr(g), ptr(inp["mask"]), ptr(inp["v0"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                       None, D(float(inp["dconst"])))
    elif op == "aad_iset_t":
        fn("aad_iset_t")(*L, ptr(g0), ptr(g), ptr(inp["mask"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "aad_iset_q":
        fn("aad_iset_q")(*L, ptr(g0), ptr(g), ptr(inp["mask"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                         D(float(inp["eps"])))
    elif op == "aad_iset_rho":
        fn("aad_iset_rho")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(inp["type"]), ptr(inp["rho"]), ptr(inp["ux"]),
                           ptr(inp["uy"]), ptr(uz), ptr(inp["tem"]), D(float(inp["eps"])))
    elif op == "ns_init":
        fn("ns_init")(*L, ptr(f0), ptr(f), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ad_init":
        fn("ad_init")(*L, ptr(g0), ptr(g), ptr(inp["tem"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ans_init":
        fn("ans_init")(*L, ptr(f0), ptr(f), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                       ptr(inp["v2"]), ptr(inp["tem"]))
    elif op == "aad_init":
        fn("aad_init")(*L, ptr(g0), ptr(g), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                       ptr(inp["v2"]), ptr(inp["tem"]))
    elif op in ("ns_collide", "ns_brinkman_collide"):
        o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz")}
        if op == "ns_collide":
            fn("ns_macro_collide")(*L, ptr(f0), ptr(f), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1), 1)
        else:
            fn("ns_macro_brinkman_collide")(*L, ptr(f0), ptr(f), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]),
                                            ptr(o["uz"]), D(0.1), ptr(inp["alpha"]), 1)
        out.update(o)
    elif op == "ad_collide_fc":
        o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        snap = np.zeros(nc * case.N)
        fn("ad_collide_fc")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]),
                            ptr(o["uz"]), ptr(inp["alpha"]), D(0.1), ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]),
                            ptr(o["qz"]), ptr(inp["diff"]), 1, ptr(snap))
        out.update(o)
        out["gsnap"] = snap.reshape(case.N, nc).T.copy()
    elif op == "ad_collide_nc":
        o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        snap = np.zeros(nc * case.N)
        fn("ad_collide_nc")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]),
                            ptr(o["uz"]), ptr(inp["alpha"]), D(0.1), ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]),
                            ptr(o["qz"]), ptr(inp["diff"]), *[D(x) for x in NC_G], D(NC_TEM0), 1, ptr(snap))
        out.update(o)
        out["gsnap"] = snap.reshape(case.N, nc).T.copy()
    elif op in ("ad_collide_bgk_fc", "ad_collide_bgk_nc"):
        o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        args = [ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1),
                ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), D(float(inp["dconst"]))]
        if op == "ad_collide_bgk_nc":
            args += [D(x) for x in NC_G] + [D(NC_TEM0)]
        fn(op)(*L, *args, 1)
        out.update(o)
    elif op == "ad_collide_he":
        beta, _, dc = he_inputs(inp)
        o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        fn("ad_collide_he")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]),
                            ptr(o["uz"]), ptr(inp["alpha"]), D(0.1), ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]),
                            ptr(o["qz"]), ptr(beta), D(dc), 1)
        out.update(o)
    elif op == "ans_collide":
        o = {k: np.zeros(case.N) for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz")}
        fn("ans_collide")(*L, ptr(f0), ptr(f), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                          ptr(o["ip"]), ptr(o["iux"]), ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]),
                          ptr(o["imz"]), D(0.1), ptr(inp["alpha"]), 1)
        out.update(o)
    elif op == "aad_collide_fc":
        o = {k: np.zeros(case.N) for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros(nc * case.N)
        fn("aad_collide_fc")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]),
                             ptr(uz), ptr(o["ip"]), ptr(o["iux"]), ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]),
                             ptr(o["imy"]), ptr(o["imz"]), ptr(inp["alpha"]), D(0.1), ptr(inp["tem"]), ptr(o["item"]),
                             ptr(o["iqx"]), ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), 1, ptr(snap))
        out.update(o)
        out["igsnap"] = snap.reshape(case.N, nc).T.copy()
    elif op == "aad_collide_nc":
        o = {k: np.zeros(case.N) for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros(nc * case.N)
        fn("aad_collide_nc")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]),
                             ptr(uz), ptr(o["ip"]), ptr(o["iux"]), ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]),
                             ptr(o["imy"]), ptr(o["imz"]), ptr(inp["alpha"]), D(0.1), ptr(inp["tem"]), ptr(o["item"]),
                             ptr(o["iqx"]), ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), *[D(x) for x in NC_G], 1,
                             ptr(snap))
        out.update(o)
        out["igsnap"] = snap.reshape(case.N, nc).T.copy()
    elif op == "aad_collide_mf":
        dx, dy, dz = mf_dirs(inp)
        o = {k: np.zeros(case.N) for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros(nc * case.N)
        fn("aad_collide_mf")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]),
                             ptr(uz), ptr(o["ip"]), ptr(o["iux"]), ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]),
                             ptr(o["imy"]), ptr(o["imz"]), ptr(inp["alpha"]), D(0.1), ptr(inp["tem"]), ptr(o["item"]),
                             ptr(o["iqx"]), ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), *[D(x) for x in NC_G],
                             ptr(dx), ptr(dy), ptr(dz), 1, ptr(snap))
        out.update(o)
        out["igsnap"] = snap.reshape(case.N, nc).T.copy()
    elif op == "aad_collide_he":
        beta, _, dc = he_inputs(inp)
        o = {k: np.zeros(case.N) for k in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        fn("aad_collide_he")(*L, ptr(f0), ptr(f), ptr(g0), ptr(g), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]),
                             ptr(uz), ptr(o["ip"]), ptr(o["iux"]), ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]),
                             ptr(o["imy"]), ptr(o["imz"]), ptr(inp["alpha"]), D(0.1), ptr(inp["tem"]), ptr(o["item"]),
                             ptr(o["iqx"]), ptr(o["iqy"]), ptr(o["iqz"]), ptr(beta), D(dc), 1)
        out.update(o)
    elif op == "sens_heat_exchange":
        _, dbds, _ = he_inputs(inp)
        d = inp["dfds0"].copy()
        fn("sens_heat_exchange")(*L, ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                                 ptr(inp["v2"]), ptr(inp["dads"]), ptr(inp["tem"]), ptr(inp["rho"]), ptr(dbds))
        out["dfds"] = d
    elif op == "sens_brinkman":
        d = inp["dfds0"].copy()
        fn("sens_brinkman")(*L, ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                            ptr(inp["v2"]), ptr(inp["dads"]))
        out["dfds"] = d
    elif op in ("sens_brinkman_diffusivity", "sens_temp_heat_source"):
        d = inp["dfds0"].copy()
        gs = np.ascontiguousarray(inp["f"].T).reshape(-1)  # snapshots in the reference layout nc*idx + c
        igs = np.ascontiguousarray(inp["g"].T).reshape(-1)
        args = [ptr(d), ptr(inp["ux"]), ptr(inp["uy"])]
        if op == "sens_brinkman_diffusivity":
            fn(op)(*L, *args, ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]), ptr(inp["dads"]),
                   ptr(inp["tem"]), ptr(inp["rho"]), ptr(inp["alpha"]), ptr(inp["uz"]), ptr(inp["uz"]), ptr(gs),
                   ptr(igs), ptr(inp["diff"]), ptr(inp["dkds"]))
        elif ND[case.kind] == 2:
            fn(op)(case.lx, case.ly, *args, ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["dads"]), ptr(inp["tem"]),
                   ptr(inp["rho"]), ptr(inp["alpha"]), ptr(inp["uz"]), ptr(gs), ptr(igs), ptr(inp["diff"]),
                   ptr(inp["dkds"]), ptr(inp["mask"]), ptr(inp["v2"]))
        else:
            # the plain Brinkman and diffusivity terms, then the face terms as the AVX build calls
            # them (the plain 3D call passes its arguments out of order; see ref_harness_avx.cpp)
            fn("sens_brinkman_diffusivity")(*L, *args, ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]),
                                            ptr(inp["dads"]), ptr(inp["tem"]), ptr(inp["rho"]), ptr(inp["alpha"]),
                                            ptr(inp["uz"]), ptr(inp["uz"]), ptr(gs), ptr(igs), ptr(inp["diff"]),
                                            ptr(inp["dkds"]))
            ref_avx().ref3avx_sens_hs_boundary(*L, ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                                               ptr(np.ascontiguousarray(inp["g"])), ptr(inp["diff"]),
                                               ptr(inp["dkds"]), ptr(inp["mask"]), ptr(inp["v2"]))
        out["dfds"] = d
    elif op == "residual":
        fr = getattr(R, case.rk + "residual")
        fr.restype = D
        a = [inp["ux"], inp["uy"], inp["uz"]]
        b = [inp["v0"], inp["v1"], inp["v2"]]
        for m in (1, 2, 3):
            out["res%d" % m] = np.array(fr(m, *[ptr(x) for x in a], *[ptr(x) for x in b], case.N))
        return out
    elif op == "normalize":
        v = inp["dfds0"].copy()
        getattr(R, case.rk + "normalize")(ptr(v), case.N)
        out["v"] = v
        return out
    elif op == "density_filter":
        v = np.zeros(case.N)
        R_ = 3.0 if ND[case.kind] == 2 else 2.0
        getattr(R, case.rk + "density_filter")(*L, D(R_), ptr(inp["tem"]), ptr(v))
        out["v"] = v
        return out
    else:
        raise KeyError(op)
    out["f"] = aos_to_soa(f0, f, nc)
    out["g"] = aos_to_soa(g0, g, nc)
    return out


def run_oracle(case, op, inp):
    O = lib()
    k = case.kind
    p = Lattice(k, case.lx, case.ly, case.lz)
    q = Lattice(case.kind_b, case.lx, case.ly, case.lz)
    p.set(inp["f"])
    q.set(inp["g"])
    F = lambda a, dt=None: case.face(a, dt)
    uz = inp["uz"] if ND[k] == 3 else None
    out = {}
    mask_f = F(inp["mask"], np.int32)
    if op == "stream":
        O.ol_stream(p.h)
    elif op == "istream":
        O.ol_istream(p.h)
    elif op in ("bc", "ibc"):
        O.ol_bc(p.h, ptr(F(inp["type"], np.int32)), int(op == "ibc"))
    elif op == "smooth":
        O.ol_smooth_corner(p.h)
    elif op == "ns_set_u":
        O.ol_ns_set_u(p.h, ptr(mask_f), ptr(F(inp["v0"])), ptr(F(inp["v1"])), ptr(F(inp["v2"])))
    elif op == "ns_set_rho":
        O.ol_ns_set_rho(p.h, ptr(mask_f), ptr(F(inp["rhob"])), ptr(F(inp["v0"])), ptr(F(inp["v1"])))
    elif op == "ans_iset_u":
        O.ol_ans_iset_u(p.h, ptr(mask_f), ptr(F(inp["v0"])), ptr(F(inp["v1"])), ptr(F(inp["v2"])),
                        D(float(inp["eps"])), 1)
    elif op == "ans_iset_rho":
        O.ol_ans_iset_rho(p.h, ptr(mask_f))
    elif op == "ad_set_t":
        O.ol_ad_set_t(q.h, ptr(mask_f), ptr(F(inp["tem"])), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ad_set_q_field":
        O.ol_ad_set_q(q.h, ptr(mask_f), ptr(F(inp["v0"])), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["diff"]),
                      D(0.0))
    elif op == "ad_set_q_const":
        O.ol_ad_set_q(q.h, ptr(mask_f), ptr(F(inp["v0"])), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), None,
                      D(float(inp["dconst"])))
    elif op == "aad_iset_t":
        O.ol_aad_iset_t(q.h, ptr(mask_f), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "aad_iset_q":
        O.ol_aad_iset_q(q.h, ptr(mask_f), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), D(float(inp["eps"])))
    elif op == "aad_iset_rho":
        O.ol_aad_iset_rho(p.h, q.h, ptr(F(inp["type"], np.int32)), ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]),
                          ptr(uz), ptr(inp["tem"]), D(float(inp["eps"])))
    elif op == "ns_init":
        O.ol_ns_init(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ad_init":
        O.ol_ad_init(q.h, ptr(inp["tem"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz))
    elif op == "ans_init":
        O.ol_ans_init(p.h, ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]),
                      ptr(inp["tem"]))
    elif op == "aad_init":
        O.ol_aad_init(q.h, ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]),
                      ptr(inp["tem"]))
    elif op in ("ns_collide", "ns_brinkman_collide"):
        o = {kk: np.zeros(case.N) for kk in ("rho", "ux", "uy", "uz")}
        if op == "ns_collide":
            O.ol_ns_macro_collide(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1), 1)
        else:
            O.ol_ns_macro_brinkman_collide(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1),
                                           ptr(inp["alpha"]), 1)
        out.update(o)
    elif op == "ad_collide_fc":
        o = {kk: np.zeros(case.N) for kk in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        snap = np.zeros((NC[case.kind_b], case.N))
        O.ol_ad_collide_fc(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), ptr(inp["alpha"]), D(0.1),
                           q.h, ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), ptr(inp["diff"]), 1,
                           ptr(snap), None)
        out.update(o)
        out["gsnap"] = snap
    elif op == "ad_collide_nc":
        o = {kk: np.zeros(case.N) for kk in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        snap = np.zeros((NC[case.kind_b], case.N))
        O.ol_ad_collide_nc(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), ptr(inp["alpha"]), D(0.1),
                           q.h, ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), ptr(inp["diff"]),
                           *[D(x) for x in NC_G], D(NC_TEM0), 1, ptr(snap))
        out.update(o)
        out["gsnap"] = snap
    elif op in ("ad_collide_bgk_fc", "ad_collide_bgk_nc"):
        o = {kk: np.zeros(case.N) for kk in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        O.ol_ad_collide_bgk(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1), q.h, ptr(o["tem"]),
                            ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), D(float(inp["dconst"])),
                            int(op == "ad_collide_bgk_nc"), *[D(x) for x in NC_G], D(NC_TEM0), 1)
        out.update(o)
    elif op == "ad_collide_he":
        beta, _, dc = he_inputs(inp)
        o = {kk: np.zeros(case.N) for kk in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
        O.ol_ad_collide_he(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), ptr(inp["alpha"]), D(0.1),
                           q.h, ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), ptr(beta), D(dc), 1)
        out.update(o)
    elif op == "ans_collide":
        o = {kk: np.zeros(case.N) for kk in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz")}
        O.ol_ans_collide(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(o["ip"]), ptr(o["iux"]),
                         ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]), ptr(o["imz"]), D(0.1),
                         ptr(inp["alpha"]), 1)
        out.update(o)
    elif op == "aad_collide_fc":
        o = {kk: np.zeros(case.N) for kk in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros((NC[case.kind_b], case.N))
        O.ol_aad_collide_fc(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(o["ip"]), ptr(o["iux"]),
                            ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]), ptr(o["imz"]),
                            ptr(inp["alpha"]), D(0.1), q.h, ptr(inp["tem"]), ptr(o["item"]), ptr(o["iqx"]),
                            ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), 1, ptr(snap))
        out.update(o)
        out["igsnap"] = snap
    elif op == "aad_collide_nc":
        o = {kk: np.zeros(case.N) for kk in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros((NC[case.kind_b], case.N))
        O.ol_aad_collide_nc(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(o["ip"]), ptr(o["iux"]),
                            ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]), ptr(o["imz"]),
                            ptr(inp["alpha"]), D(0.1), q.h, ptr(inp["tem"]), ptr(o["item"]), ptr(o["iqx"]),
                            ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), *[D(x) for x in NC_G], 1, ptr(snap))
        out.update(o)
        out["igsnap"] = snap
    elif op == "aad_collide_mf":
        dx, dy, dz = mf_dirs(inp)
        o = {kk: np.zeros(case.N) for kk in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        snap = np.zeros((NC[case.kind_b], case.N))
        O.ol_aad_collide_mf(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(o["ip"]), ptr(o["iux"]),
                            ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]), ptr(o["imz"]),
                            ptr(inp["alpha"]), D(0.1), q.h, ptr(inp["tem"]), ptr(o["item"]), ptr(o["iqx"]),
                            ptr(o["iqy"]), ptr(o["iqz"]), ptr(inp["diff"]), *[D(x) for x in NC_G], ptr(dx), ptr(dy),
                            ptr(dz) if ND[case.kind] == 3 else None, 1, ptr(snap))
        out.update(o)
        out["igsnap"] = snap
    elif op == "aad_collide_he":
        beta, _, dc = he_inputs(inp)
        o = {kk: np.zeros(case.N) for kk in ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")}
        O.ol_aad_collide_he(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(o["ip"]), ptr(o["iux"]),
                            ptr(o["iuy"]), ptr(o["iuz"]), ptr(o["imx"]), ptr(o["imy"]), ptr(o["imz"]),
                            ptr(inp["alpha"]), D(0.1), q.h, ptr(inp["tem"]), ptr(o["item"]), ptr(o["iqx"]),
                            ptr(o["iqy"]), ptr(o["iqz"]), ptr(beta), D(dc), 1)
        out.update(o)
    elif op == "sens_heat_exchange":
        _, dbds, _ = he_inputs(inp)
        d = inp["dfds0"].copy()
        O.ol_sens_heat_exchange(p.h, ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                                ptr(inp["v2"]), ptr(inp["dads"]), ptr(inp["tem"]), ptr(inp["rho"]), ptr(dbds))
        out["dfds"] = d
    elif op == "sens_brinkman":
        d = inp["dfds0"].copy()
        O.ol_sens_brinkman(p.h, ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]),
                           ptr(inp["v2"]), ptr(inp["dads"]))
        out["dfds"] = d
    elif op in ("sens_brinkman_diffusivity", "sens_temp_heat_source"):
        d = inp["dfds0"].copy()
        gs = np.ascontiguousarray(inp["f"])
        igs = np.ascontiguousarray(inp["g"])
        common = [ptr(d), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz), ptr(inp["v0"]), ptr(inp["v1"]), ptr(inp["v2"]),
                  ptr(inp["dads"]), ptr(inp["tem"]), ptr(inp["rho"]), ptr(inp["alpha"]), ptr(inp["uz"]),
                  ptr(inp["uz"]), ptr(gs), ptr(igs), ptr(inp["diff"]), ptr(inp["dkds"])]
        if op == "sens_brinkman_diffusivity":
            O.ol_sens_brinkman_diffusivity(q.h, *common)
        else:
            O.ol_sens_temp_heat_source(q.h, *common, ptr(mask_f), ptr(F(inp["v2"])))
        out["dfds"] = d
    elif op == "residual":
        a = [inp["ux"], inp["uy"], inp["uz"]]
        b = [inp["v0"], inp["v1"], inp["v2"]]
        for m in (1, 2, 3):
            out["res%d" % m] = np.array(O.ol_residual(m, *[ptr(x) for x in a], *[ptr(x) for x in b], case.N))
        return out
    elif op == "normalize":
        v = inp["dfds0"].copy()
        O.ol_normalize(ptr(v), case.N)
        out["v"] = v
        return out
    elif op == "density_filter":
        v = np.zeros(case.N)
        R_ = 3.0 if ND[k] == 2 else 2.0
        O.ol_density_filter(case.lx, case.ly, case.lz, D(R_), ptr(inp["tem"]), ptr(v))
        out["v"] = v
        return out
    else:
        raise KeyError(op)
    out["f"] = p.get()
    out["g"] = q.get()
    return out


OPS_2D = {
    "stream": "d2q9.h:282-291", "istream": "d2q9.h:414-423", "bc": "d2q9.h:429-501",
    "ibc": "d2q9.h:503-575", "smooth": "d2q9.h:577-587",
    "ns_set_u": "navierstokes.h:92-146", "ns_set_rho": "navierstokes.h:261-315",
    "ans_iset_u": "adjointnavierstokes.h:97-143", "ans_iset_rho": "adjointnavierstokes.h:258-302",
    "ad_set_t": "advection.h:99-145", "ad_set_q_field": "advection.h:385-431", "ad_set_q_const": "advection.h:242-288",
    "aad_iset_t": "adjointadvection.h:154-198", "aad_iset_q": "adjointadvection.h:304-364",
    "aad_iset_rho": "adjointadvection.h:488-574",
    "ns_init": "navierstokes.h:550-559", "ad_init": "advection.h:1048-1057", "ans_init": "adjointnavierstokes.h:474-484",
    "aad_init": "adjointadvection.h:1359-1368",
    "ns_collide": "navierstokes.h:432-455", "ns_brinkman_collide": "navierstokes.h:487-514",
    "ad_collide_fc": "advection.h:814-867", "ans_collide": "adjointnavierstokes.h:398-431",
    "aad_collide_fc": "adjointadvection.h:997-1050",
    "ad_collide_nc": "advection.h:930-984,73-79", "aad_collide_nc": "adjointadvection.h:1113-1168,114-122",
    "ad_collide_he": "advection.h:716-760,89-95", "aad_collide_he": "adjointadvection.h:894-940,106-112",
    "ad_collide_bgk_fc": "advection.h:528-569", "ad_collide_bgk_nc": "advection.h:616-662,73-79",
    "aad_collide_mf": "adjointadvection.h:1232-1291,134-141",
    "sens_heat_exchange": "adjointadvection.h:1445-1450",
    "sens_brinkman": "adjointnavierstokes.h:542-546", "sens_brinkman_diffusivity": "adjointadvection.h:1461-1475",
    "sens_temp_heat_source": "adjointadvection.h:1497-1517,755-801",
    "residual": "residual.h:9-51", "normalize": "normalize.h:9-24", "density_filter": "densityfilter.h:390-420,512-515",
}

# Functions whose arithmetic calls glibc pow(): value equality is not expected on other libms.
POW_OPS = {"sens_brinkman_diffusivity", "sens_temp_heat_source", "residual", "density_filter"}

CASES_2D = [Case("d2q9_12x10", D2Q9, (12, 10), 11, "D2Q9"), Case("d2q9_7x9", D2Q9, (7, 9), 12, "D2Q9")]

CASES_3D = [Case("d3q15_5x4x6", D3Q15, (5, 4, 6), 21, "D3Q15"), Case("d3q15_4x6x5", D3Q15, (4, 6, 5), 22, "D3Q15")]

# D3Q15.  AAD::iBoundaryConditionSetRho has no 3D case: the reference's D3Q15 overload does not
# compile (adjointadvection.h:583 names an undeclared `_bctype`).  The 3D boundary term of
# SensitivityTemperatureAtHeatSource is taken from the reference's AVX build (oracle/_ref/libref_avx.so).
OPS_3D = {
    "stream": "d3q15.h:602-613", "istream": "d3q15.h:965-976", "bc": "d3q15.h:982-1110",
    "ibc": "d3q15.h:1112-1239", "smooth": "d3q15.h:199-220,1241-1303",
    "ns_set_u": "navierstokes.h:150-257", "ns_set_rho": "navierstokes.h:319-426",
    "ans_iset_u": "adjointnavierstokes.h:147-254", "ans_iset_rho": "adjointnavierstokes.h:306-392",
    "ad_set_t": "advection.h:149-238", "ad_set_q_field": "advection.h:435-524", "ad_set_q_const": "advection.h:292-381",
    "aad_iset_t": "adjointadvection.h:202-300", "aad_iset_q": "adjointadvection.h:368-484",
    "ns_init": "navierstokes.h:563-572", "ad_init": "advection.h:1061-1070", "ans_init": "adjointnavierstokes.h:488-498",
    "aad_init": "adjointadvection.h:1372-1381",
    "ns_collide": "navierstokes.h:459-483", "ns_brinkman_collide": "navierstokes.h:518-546",
    "ad_collide_fc": "advection.h:871-926", "ans_collide": "adjointnavierstokes.h:435-470",
    "aad_collide_fc": "adjointadvection.h:1054-1109",
    "ad_collide_nc": "advection.h:988-1044,81-87", "aad_collide_nc": "adjointadvection.h:1172-1230,124-132",
    "ad_collide_he": "advection.h:764-810,89-95", "aad_collide_he": "adjointadvection.h:944-993,106-112",
    "ad_collide_bgk_fc": "advection.h:571-614", "ad_collide_bgk_nc": "advection.h:664-712,81-87",
    "aad_collide_mf": "adjointadvection.h:1293-1358,143-150",
    "sens_heat_exchange": "adjointadvection.h:1452-1457",
    "sens_brinkman": "adjointnavierstokes.h:550-554", "sens_brinkman_diffusivity": "adjointadvection.h:1478-1493",
    "sens_temp_heat_source": "adjointadvection.h:1519-1543,805-888 (argument order of equation_avx/adjointadvection_avx.h:1507-1512)",
    "residual": "residual.h:39-51", "normalize": "normalize.h:9-24", "density_filter": "densityfilter.h:390-420,512-515",
}


# D3Q27 (no reference counterpart) coupled with D3Q7 advection-diffusion: every 3D function, the
# expected values from the oracle (tests/test_gpu_d3q27.py), which tests/test_oracle_d3q27.py pins.
CASES_D3Q27 = [Case("d3q27_5x4x6", D3Q27, (5, 4, 6), 31, "D3Q27", kind_b=D3Q7),
               Case("d3q27_4x6x5", D3Q27, (4, 6, 5), 32, "D3Q27", kind_b=D3Q7)]


def ops_for(case):
    return list(OPS_2D) if ND[case.kind] == 2 else list(OPS_3D)
