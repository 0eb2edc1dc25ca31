"""GPU parity of the sensitivity terms the reference fixtures do not cover: the volumetric heat-source
term (lbm_sensitivity_heat_source) on every advection-diffusion lattice, and the heat-source boundary
term on D2Q5 / D3Q7 and in both compat modes on D3Q15, against the CPU oracle (whose forms
tests/test_sensitivity_fd.py pins by finite differences)."""
import ctypes

import numpy as np
import pytest

from _oracle import D2Q5, D2Q9, D3Q7, D3Q15, NC, ND, Lattice, face_nodes, lib, ptr

pytestmark = pytest.mark.gpu
D = ctypes.c_double


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _inputs(kind, seed):
    lx, ly, lz = (7, 6, 1) if ND[kind] == 2 else (5, 6, 4)
    N = lx * ly * lz
    rng = np.random.default_rng(seed)
    d = {k: 0.05 * rng.uniform(-1, 1, N) for k in ("ux", "uy", "uz", "imx", "imy", "imz")}
    d.update({k: rng.standard_normal(N) for k in ("item", "iqx", "iqy", "iqz", "tem")})
    d["diff"] = rng.uniform(0.02, 0.2, N)
    d["dkds"] = rng.uniform(-0.2, -0.05, N)
    d["dads"] = rng.uniform(-50, -1, N)
    d["dhds"] = np.full(N, -1e-3)
    d["dfds0"] = rng.standard_normal(N)
    d["g"] = np.abs(rng.standard_normal((NC[kind], N)))
    d["ig"] = rng.standard_normal((NC[kind], N))
    fn = face_nodes(kind, lx, ly, lz)
    d["mask_f"] = (rng.uniform(size=fn.size) < 0.6).astype(np.int32)
    d["qn_f"] = rng.uniform(0.5, 1.5, fn.size) * 1e-2
    return (lx, ly, lz), d


@pytest.mark.parametrize("kind", [D2Q9, D2Q5, D3Q15, D3Q7], ids=["D2Q9", "D2Q5", "D3Q15", "D3Q7"])
def test_heat_source_sensitivity_matches_oracle(ctx, kind):
    import panslbm2_amd as pa

    print("heat source sensitivity", kind, flush=True)
    (lx, ly, lz), d = _inputs(kind, 7 + kind)
    N = lx * ly * lz
    d3 = ND[kind] == 3
    q = Lattice(kind, lx, ly, lz)
    want = d["dfds0"].copy()
    P = lambda a: ptr(a) if d3 else None
    lib().ol_sens_heat_source(q.h, ptr(want), ptr(d["ux"]), ptr(d["uy"]), P(d["uz"]), ptr(d["item"]), ptr(d["iqx"]),
                              ptr(d["iqy"]), P(d["iqz"]), ptr(d["diff"]), ptr(d["dhds"]))
    gq = pa.Lattice(ctx, kind, lx, ly, lz)
    a = {k: ctx.array((N,), data=d[k]) for k in ("ux", "uy", "uz", "item", "iqx", "iqy", "iqz", "diff", "dhds")}
    dfds = ctx.array((N,), data=d["dfds0"])
    G = lambda k: a[k].ptr if (d3 or k not in ("uz", "iqz")) else None
    pa.check(ctx.lib.lbm_sensitivity_heat_source(gq.h, dfds.ptr, G("ux"), G("uy"), G("uz"), G("item"), G("iqx"),
                                                 G("iqy"), G("iqz"), G("diff"), G("dhds")))
    np.testing.assert_array_equal(dfds.download(), want)


@pytest.mark.parametrize("kind,compat", [(D2Q5, 1), (D3Q7, 1), (D3Q15, 0), (D3Q15, 1)],
                         ids=["D2Q5", "D3Q7", "D3Q15-c0", "D3Q15-c1"])
def test_heat_source_boundary_matches_oracle(ctx, kind, compat):
    import panslbm2_amd as pa

    print("heat-source boundary sensitivity", kind, compat, flush=True)
    (lx, ly, lz), d = _inputs(kind, 30 + kind + compat)
    N = lx * ly * lz
    d3 = ND[kind] == 3
    q = Lattice(kind, lx, ly, lz)
    P = lambda k: ptr(d[k]) if d3 else None
    want = d["dfds0"].copy()
    lib().ol_sens_temp_heat_source_c(q.h, ptr(want), ptr(d["ux"]), ptr(d["uy"]), P("uz"), ptr(d["imx"]),
                                     ptr(d["imy"]), P("imz"), ptr(d["dads"]), ptr(d["tem"]), ptr(d["item"]),
                                     ptr(d["iqx"]), ptr(d["iqy"]), P("iqz"), ptr(d["g"]), ptr(d["ig"]), ptr(d["diff"]),
                                     ptr(d["dkds"]), ptr(d["mask_f"]), ptr(d["qn_f"]), compat)
    gq = pa.Lattice(ctx, kind, lx, ly, lz)
    keys = ("ux", "uy", "uz", "imx", "imy", "imz", "dads", "tem", "item", "iqx", "iqy", "iqz", "diff", "dkds")
    a = {k: ctx.array((N,), data=d[k]) for k in keys}
    g = ctx.array(d["g"].shape, data=d["g"])
    ig = ctx.array(d["ig"].shape, data=d["ig"])
    dfds = ctx.array((N,), data=d["dfds0"])
    G = lambda k: a[k].ptr if (d3 or k not in ("uz", "imz", "iqz")) else None
    hp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    pa.check(ctx.lib.lbm_sensitivity_temperature_at_heat_source_c(
        gq.h, dfds.ptr, G("ux"), G("uy"), G("uz"), G("imx"), G("imy"), G("imz"), G("dads"), G("tem"), G("item"),
        G("iqx"), G("iqy"), G("iqz"), g.ptr, ig.ptr, G("diff"), G("dkds"), hp(d["mask_f"]), hp(d["qn_f"]), compat))
    # the diffusivity term squares with x*x on the device and pow(x, 2.0) in the oracle: both are the
    # correctly rounded square, but the sums differ in association, hence not bit-for-bit
    np.testing.assert_allclose(dfds.download(), want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
