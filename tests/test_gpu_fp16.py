"""FP16 population storage (LBM_F16) against the FP64 reference fixtures and the FP64 path.

LBM_F16 stores the same deviations as LBM_F32 in IEEE binary16 (11 significant bits) and collides
them in FP32.  Bars:
  * one call of every population function on inputs uploaded with zero shift (each population
    rounded to binary16 once): max |FP16 - reference| <= TOL_CALL * max |reference| per output
    array.  binary16 rounds each population (size ~w) to 2^-11 = 4.9e-4 of itself, and a velocity
    is a difference of populations 10-20x its size, so the velocities carry up to ~1e-2 (7.7e-3
    measured); in a run, where an NS lattice stores deviations from the rest state, the same
    rounding is ~20x smaller;
  * streaming and lattice boundary copies move binary16 values unchanged: exact against the
    reference rounded to binary16;
  * an equilibrium initial condition (NS: deviations from the rest state) round-trips through
    FP64 exactly;
  * a steady heat-sink evaluation run in FP16 until its residual reaches the FP16 floor (1e-3),
    finished in FP64, both solved to 1e-9 so that only precision separates them: the objective
    to 1e-5 and the design sensitivity to 1e-4 of its largest magnitude, as for FP32.
"""
import os

import numpy as np
import pytest

import refcases

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL_CALL = 1e-2
COPY_OPS = {"stream", "istream", "bc", "ibc"}
SKIP = ("residual", "normalize", "density_filter", "sens_brinkman", "sens_brinkman_diffusivity",
        "sens_temp_heat_source", "sens_heat_exchange")


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _relerr(got, want):
    got, want = np.asarray(got, dtype=np.float64), np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape
    assert np.all(np.isfinite(got)), "non-finite values"
    scale = max(float(np.max(np.abs(want))), 1e-300)
    return float(np.max(np.abs(got - want))) / scale


def _cases():
    return [(c, op) for c in refcases.CASES_2D + refcases.CASES_3D for op in refcases.ops_for(c) if op not in SKIP]


@pytest.mark.parametrize("case,op", _cases(), ids=lambda x: getattr(x, "name", x))
def test_fp16_function_matches_reference(ctx, case, op):
    import panslbm2_amd as pa
    from gpu_cases import run_gpu

    print("fp16 case", case.name, op, flush=True)
    blob = np.load(os.path.join(GOLDEN, case.name + ".npz"))
    inp = {k[3:]: blob[k] for k in blob.files if k.startswith("in/")}
    pre = "out/%s/" % op
    want = {k[len(pre):]: blob[k] for k in blob.files if k.startswith(pre)}
    got = run_gpu(ctx, case, op, inp, prec=pa.F16)
    for k, v in want.items():
        if op in COPY_OPS and k in ("f", "g"):
            np.testing.assert_array_equal(np.asarray(got[k]), np.asarray(v).astype(np.float16).astype(np.float64),
                                          err_msg="%s %s %s" % (case.name, op, k))
            continue
        err = _relerr(got[k], v)
        print("  %-8s rel err %.3g" % (k, err), flush=True)
        assert err <= TOL_CALL, "%s %s %s: relative error %.3g > %.1g" % (case.name, op, k, err, TOL_CALL)


@pytest.mark.parametrize("eq,kind", [("ns", "D2Q9"), ("ns", "D3Q15"), ("ns", "D3Q27"), ("ad", "D2Q5"),
                                     ("ad", "D3Q7")])
def test_fp16_storage_round_trip(ctx, eq, kind):
    import panslbm2_amd as pa

    k = getattr(pa, kind)
    dims = (7, 5, 4) if pa.ND[k] == 3 else (9, 6, 1)
    N = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(4)
    rho, ux, uy, uz = 1 + 0.02 * rng.standard_normal(N), *(0.05 * rng.standard_normal((3, N)))
    tem = rng.uniform(0.0, 1.5, N)
    a = {n: ctx.array((N,), data=v) for n, v in dict(rho=rho, ux=ux, uy=uy, uz=uz, tem=tem).items()}
    uzp = a["uz"].ptr if pa.ND[k] == 3 else None
    out = {}
    for prec in (pa.F64, pa.F16):
        lat = pa.Lattice(ctx, k, *dims, prec=prec)
        if eq == "ns":
            pa.check(ctx.lib.lbm_ns_initial_condition(lat.h, a["rho"].ptr, a["ux"].ptr, a["uy"].ptr, uzp))
        else:
            pa.check(ctx.lib.lbm_ad_initial_condition(lat.h, a["tem"].ptr, a["ux"].ptr, a["uy"].ptr, uzp))
        out[prec] = lat.download()
        if prec == pa.F16:
            pa.check(ctx.lib.lbm_lattice_set_precision(lat.h, pa.F64))
            np.testing.assert_array_equal(lat.download(), out[prec])
            pa.check(ctx.lib.lbm_lattice_set_precision(lat.h, pa.F16))
            np.testing.assert_array_equal(lat.download(), out[prec])
        lat.close()
    err = _relerr(out[pa.F16], out[pa.F64])
    print("init rel err %.3g" % err, flush=True)
    # NS: deviations of size ~0.15 w rounded to 2^-11; AD: the values themselves
    assert err <= (1e-4 if eq == "ns" else 5e-4)


def test_fp16_heatsink_mixed_precision(ctx):
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink2d import Heatsink2D

    out = {}
    for prec in (pa.F64, pa.F16):
        hs = Heatsink2D(ctx, lx=48, ly=96, ad_kind=pa.D2Q5, seed=3, filter_radius=3.0, max_steps=200000, prec=prec,
                        eps=1e-9)
        r = hs.evaluate()
        out[prec] = (r, hs.d["dfds"].download())
        print("prec", prec, "steps", r["fwd_steps"], r["adj_steps"], "in fp16", r.get("fp32_steps"), "objective",
              r["objective"], flush=True)
        assert max(r["fwd_residual"]) < 1e-9 and max(r["adj_residual"]) < 1e-9
        hs.close()
    (r64, s64), (r16, s16) = out[pa.F64], out[pa.F16]
    es = _relerr(s16, s64)
    eo = abs(r16["objective"] - r64["objective"]) / abs(r64["objective"])
    print("dfds rel err %.3g, objective rel err %.3g" % (es, eo), flush=True)
    assert r16["fp32_steps"][0] > 0  # the forward solve did run in FP16 first
    assert es <= 1e-4
    assert eo <= 1e-5
