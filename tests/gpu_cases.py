"""Run one reference function of tests/refcases.py on the HIP library (eager C-ABI calls)."""
import ctypes

import numpy as np

import panslbm2_amd as pa
from panslbm2_amd import D, check, hptr

from _oracle import NC, ND
from refcases import NC_G, NC_TEM0, he_inputs, mf_dirs


def run_gpu(ctx, case, op, inp, prec=pa.F64):
    L = ctx.lib
    k = case.kind
    N = case.N
    p = pa.Lattice(ctx, k, case.lx, case.ly, case.lz, prec=prec)
    q = pa.Lattice(ctx, case.kind_b, case.lx, case.ly, case.lz, prec=prec)
    p.upload(inp["f"])
    q.upload(inp["g"])
    F = lambda a, dt=np.float64: pa.to_face(k, case.lx, case.ly, case.lz, a, dt)
    dev = {}

    def d(name, arr=None):
        if name not in dev:
            a = inp[name] if arr is None else arr
            dev[name] = ctx.array(np.asarray(a).shape, data=np.asarray(a, dtype=np.float64))
        return dev[name].ptr

    uzp = d("uz") if ND[k] == 3 else None
    mask_f = F(inp["mask"], np.int32)
    out = {}

    def outs(names, shape=None):
        for n in names:
            dev["o_" + n] = ctx.array(shape or (N,), fill=0.0)
        return [dev["o_" + n].ptr for n in names]

    if op == "stream":
        check(L.lbm_stream(p.h))
    elif op == "istream":
        check(L.lbm_istream(p.h))
    elif op == "bc":
        check(L.lbm_bc(p.h, hptr(F(inp["type"], np.int32))))
    elif op == "ibc":
        check(L.lbm_ibc(p.h, hptr(F(inp["type"], np.int32))))
    elif op == "smooth":
        check(L.lbm_smooth_corner(p.h))
    elif op == "ns_set_u":
        check(L.lbm_ns_set_u(p.h, hptr(mask_f), hptr(F(inp["v0"])), hptr(F(inp["v1"])), hptr(F(inp["v2"]))))
    elif op == "ns_set_rho":
        check(L.lbm_ns_set_rho(p.h, hptr(mask_f), hptr(F(inp["rhob"])), hptr(F(inp["v0"])), hptr(F(inp["v1"]))))
    elif op == "ans_iset_u":
        check(L.lbm_ans_iset_u(p.h, hptr(mask_f), hptr(F(inp["v0"])), hptr(F(inp["v1"])), hptr(F(inp["v2"])),
                               D(float(inp["eps"])), 1))
    elif op == "ans_iset_rho":
        check(L.lbm_ans_iset_rho(p.h, hptr(mask_f)))
    elif op == "ad_set_t":
        check(L.lbm_ad_set_t(q.h, hptr(mask_f), hptr(F(inp["tem"])), d("ux"), d("uy"), uzp))
    elif op == "ad_set_q_field":
        check(L.lbm_ad_set_q(q.h, hptr(mask_f), hptr(F(inp["v0"])), d("ux"), d("uy"), uzp, d("diff"), D(0.0)))
    elif op == "ad_set_q_const":
        check(L.lbm_ad_set_q(q.h, hptr(mask_f), hptr(F(inp["v0"])), d("ux"), d("uy"), uzp, None,
                             D(float(inp["dconst"]))))
    elif op == "aad_iset_t":
        check(L.lbm_aad_iset_t(q.h, hptr(mask_f), d("ux"), d("uy"), uzp))
    elif op == "aad_iset_q":
        check(L.lbm_aad_iset_q(q.h, hptr(mask_f), d("ux"), d("uy"), uzp, D(float(inp["eps"]))))
    elif op == "aad_iset_rho":
        check(L.lbm_aad_iset_rho(p.h, q.h, hptr(F(inp["type"], np.int32)), d("rho"), d("ux"), d("uy"), uzp, d("tem"),
                                 D(float(inp["eps"]))))
    elif op == "ns_init":
        check(L.lbm_ns_initial_condition(p.h, d("rho"), d("ux"), d("uy"), uzp))
    elif op == "ad_init":
        check(L.lbm_ad_initial_condition(q.h, d("tem"), d("ux"), d("uy"), uzp))
    elif op == "ans_init":
        check(L.lbm_ans_initial_condition(p.h, d("ux"), d("uy"), uzp, d("v0"), d("v1"), d("v2"), d("tem")))
    elif op == "aad_init":
        check(L.lbm_aad_initial_condition(q.h, d("ux"), d("uy"), uzp, d("v0"), d("v1"), d("v2"), d("tem")))
    elif op in ("ns_collide", "ns_brinkman_collide"):
        names = ("rho", "ux", "uy", "uz")
        o = outs(names)
        if op == "ns_collide":
            check(L.lbm_ns_macro_collide(p.h, *o, D(0.1), 1))
        else:
            check(L.lbm_ns_macro_brinkman_collide(p.h, *o, D(0.1), d("alpha"), 1))
        for n in names:
            out[n] = dev["o_" + n].download()
    elif op == "ad_collide_fc":
        names = ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")
        o = outs(names)
        snap = outs(["gsnap"], (NC[case.kind_b], N))[0]
        check(L.lbm_ad_macro_brinkman_collide_force_convection(p.h, *o[:4], d("alpha"), D(0.1), q.h, *o[4:],
                                                               d("diff"), 1, snap, None))
        for n in names + ("gsnap",):
            out[n] = dev["o_" + n].download()
    elif op == "ad_collide_nc":
        names = ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")
        o = outs(names)
        snap = outs(["gsnap"], (NC[case.kind_b], N))[0]
        check(L.lbm_ad_macro_brinkman_collide_natural_convection(p.h, *o[:4], d("alpha"), D(0.1), q.h, *o[4:],
                                                                 d("diff"), *[D(x) for x in NC_G], D(NC_TEM0), 1,
                                                                 snap))
        for n in names + ("gsnap",):
            out[n] = dev["o_" + n].download()
    elif op in ("ad_collide_bgk_fc", "ad_collide_bgk_nc"):
        names = ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")
        o = outs(names)
        if op == "ad_collide_bgk_fc":
            check(L.lbm_ad_macro_collide_force_convection(p.h, *o[:4], D(0.1), q.h, *o[4:], D(float(inp["dconst"])),
                                                          1))
        else:
            check(L.lbm_ad_macro_collide_natural_convection(p.h, *o[:4], D(0.1), q.h, *o[4:],
                                                            D(float(inp["dconst"])), *[D(x) for x in NC_G],
                                                            D(NC_TEM0), 1))
        for n in names:
            out[n] = dev["o_" + n].download()
    elif op == "ad_collide_he":
        beta, _, dc = he_inputs(inp)
        names = ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")
        o = outs(names)
        check(L.lbm_ad_macro_brinkman_collide_heat_exchange(p.h, *o[:4], d("alpha"), D(0.1), q.h, *o[4:],
                                                            d("beta", beta), D(dc), 1))
        for n in names:
            out[n] = dev["o_" + n].download()
    elif op == "ans_collide":
        names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz")
        o = outs(names)
        check(L.lbm_ans_macro_brinkman_collide(p.h, d("rho"), d("ux"), d("uy"), uzp, *o, D(0.1), d("alpha"), 1))
        for n in names:
            out[n] = dev["o_" + n].download()
    elif op == "aad_collide_fc":
        names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")
        o = outs(names)
        snap = outs(["igsnap"], (NC[case.kind_b], N))[0]
        check(L.lbm_aad_macro_brinkman_collide_force_convection(p.h, d("rho"), d("ux"), d("uy"), uzp, *o[:7],
                                                               d("alpha"), D(0.1), q.h, d("tem"), *o[7:], d("diff"), 1,
                                                               snap))
        for n in names + ("igsnap",):
            out[n] = dev["o_" + n].download()
    elif op == "aad_collide_nc":
        names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")
        o = outs(names)
        snap = outs(["igsnap"], (NC[case.kind_b], N))[0]
        check(L.lbm_aad_macro_brinkman_collide_natural_convection(p.h, d("rho"), d("ux"), d("uy"), uzp, *o[:7],
                                                                  d("alpha"), D(0.1), q.h, d("tem"), *o[7:],
                                                                  d("diff"), *[D(x) for x in NC_G], 1, snap))
        for n in names + ("igsnap",):
            out[n] = dev["o_" + n].download()
    elif op == "aad_collide_mf":
        dx, dy, dz = mf_dirs(inp)
        names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")
        o = outs(names)
        snap = outs(["igsnap"], (NC[case.kind_b], N))[0]
        check(L.lbm_aad_macro_brinkman_collide_natural_convection_mass_flow(
            p.h, d("rho"), d("ux"), d("uy"), uzp, *o[:7], d("alpha"), D(0.1), q.h, d("tem"), *o[7:], d("diff"),
            *[D(x) for x in NC_G], d("dirx", dx), d("diry", dy), d("dirz", dz) if uzp else None, 1, snap))
        for n in names + ("igsnap",):
            out[n] = dev["o_" + n].download()
    elif op == "aad_collide_he":
        beta, _, dc = he_inputs(inp)
        names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")
        o = outs(names)
        check(L.lbm_aad_macro_brinkman_collide_heat_exchange(p.h, d("rho"), d("ux"), d("uy"), uzp, *o[:7],
                                                             d("alpha"), D(0.1), q.h, d("tem"), *o[7:],
                                                             d("beta", beta), D(dc), 1))
        for n in names:
            out[n] = dev["o_" + n].download()
    elif op == "sens_heat_exchange":
        _, dbds, _ = he_inputs(inp)
        dfds = ctx.array((N,), data=inp["dfds0"])
        check(L.lbm_sensitivity_heat_exchange(q.h, dfds.ptr, d("ux"), d("uy"), d("uz"), d("v0"), d("v1"), d("v2"),
                                              d("dads"), d("tem"), d("rho"), d("dbds", dbds)))
        out["dfds"] = dfds.download()
    elif op == "sens_brinkman":
        dfds = ctx.array((N,), data=inp["dfds0"])
        check(L.lbm_sensitivity_brinkman(p.h, dfds.ptr, d("ux"), d("uy"), d("uz"), d("v0"), d("v1"), d("v2"),
                                         d("dads")))
        out["dfds"] = dfds.download()
    elif op in ("sens_brinkman_diffusivity", "sens_temp_heat_source"):
        dfds = ctx.array((N,), data=inp["dfds0"])
        g = ctx.array(inp["f"].shape, data=inp["f"])
        ig = ctx.array(inp["g"].shape, data=inp["g"])
        common = [dfds.ptr, d("ux"), d("uy"), d("uz"), d("v0"), d("v1"), d("v2"), d("dads"), d("tem"), d("rho"),
                  d("alpha"), d("uz"), d("uz"), g.ptr, ig.ptr, d("diff"), d("dkds")]
        if op == "sens_brinkman_diffusivity":
            check(L.lbm_sensitivity_brinkman_diffusivity(q.h, *common))
        else:
            check(L.lbm_sensitivity_temperature_at_heat_source(q.h, *common, hptr(mask_f), hptr(F(inp["v2"]))))
        out["dfds"] = dfds.download()
    elif op == "residual":
        a = [d("ux"), d("uy"), d("uz")]
        b = [d("v0"), d("v1"), d("v2")]
        for m in (1, 2, 3):
            r = ctypes.c_double()
            check(L.lbm_residual(ctx.h, m, *(a[:m] + [None] * (3 - m)), *(b[:m] + [None] * (3 - m)), N,
                                 ctypes.byref(r)))
            out["res%d" % m] = np.array(r.value)
        return out
    elif op == "normalize":
        v = ctx.array((N,), data=inp["dfds0"])
        check(L.lbm_normalize(ctx.h, v.ptr, N))
        out["v"] = v.download()
        return out
    elif op == "density_filter":
        v = ctx.array((N,), fill=0.0)
        R_ = 3.0 if ND[k] == 2 else 2.0
        check(L.lbm_density_filter(ctx.h, case.lx, case.ly, case.lz, D(R_), d("tem"), v.ptr))
        out["v"] = v.download()
        return out
    else:
        raise KeyError(op)
    out["f"] = p.download()
    out["g"] = q.download()
    return out
