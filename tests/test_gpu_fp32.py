"""FP32 population storage (LBM_F32) against the FP64 reference fixtures and the FP64 path.

LBM_F32 stores the populations in FP32 (an NS lattice's as deviations from the rest state) and
collides them in FP32; the bar is a stated tolerance:
  * one call of every population function (stream, boundaries, smoothing, collides, initial
    conditions) on inputs uploaded with zero shift (each population rounded to FP32 once):
    max |FP32 - reference| <= TOL_CALL * max |reference| per output array.  A velocity is a
    difference of populations ~10x its size, so input rounding alone gives ~1e-6 of it.
  * whole driver sequences (a few hundred steps): TOL_SEQ relative to each array's max.
  * the steady design sensitivity of the heat-sink workload at the 1e-6 residual, solved in FP32
    and finished in FP64 once the FP32 residual stops falling (the heat-flux residual has an FP32
    floor near 1e-5: the advective part T u of the AD first moment exceeds the diffusive flux q by
    the Peclet number, and u carries FP32 rounding), against the FP64 path (itself value-identical
    to the reference): max |d dfds| <= 1e-4 * max |dfds_64|, the objective to 1e-5 relative.
Streaming and lattice boundary copies move FP32 values unchanged, so those are exact against the
FP64 result rounded to FP32.
"""
import os

import numpy as np
import pytest

import refcases
import sequences

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL_CALL = 1e-5
TOL_SEQ = 1e-4
COPY_OPS = {"stream", "istream", "bc", "ibc"}


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
    return [(c, op) for c in refcases.CASES_2D + refcases.CASES_3D for op in refcases.ops_for(c)
            if op not in ("residual", "normalize", "density_filter", "sens_brinkman", "sens_brinkman_diffusivity",
                          "sens_temp_heat_source")]


@pytest.mark.parametrize("case,op", _cases(), ids=lambda x: getattr(x, "name", x))
def test_fp32_function_matches_reference(ctx, case, op):
    import panslbm2_amd as pa
    from gpu_cases import run_gpu

    print("fp32 case", case.name, op, flush=True)
    blob = np.load(os.path.join(GOLDEN, case.name + ".npz"))
    inp = {k[3:]: blob[k] for k in blob.files if k.startswith("in/")}
    pre = "out/%s/" % op
    want = {k[len(pre):]: blob[k] for k in blob.files if k.startswith(pre)}
    got = run_gpu(ctx, case, op, inp, prec=pa.F32)
    for k, v in want.items():
        if op in COPY_OPS and k in ("f", "g"):
            # the inputs rounded to FP32 are moved unchanged
            np.testing.assert_array_equal(np.asarray(got[k]), np.asarray(v).astype(np.float32).astype(np.float64),
                                          err_msg="%s %s %s" % (case.name, op, k))
            continue
        err = _relerr(got[k], v)
        print("  %-8s rel err %.3g" % (k, err), flush=True)
        assert err <= TOL_CALL, "%s %s %s: relative error %.3g > %.1g" % (case.name, op, k, err, TOL_CALL)


@pytest.mark.parametrize("fused", [False, True], ids=["eager", "fused"])
@pytest.mark.parametrize("seq", sequences.SEQUENCES, ids=lambda s: s.name)
def test_fp32_sequence_matches_reference(ctx, seq, fused):
    import panslbm2_amd as pa
    from gpu_backend import FusedGpuBackend, GpuBackend

    print("fp32 sequence", seq.name, "fused" if fused else "eager", flush=True)
    b = FusedGpuBackend(ctx, pa.F32) if fused else GpuBackend(ctx, pa.F32)
    got = seq.run(b)
    blob = np.load(os.path.join(GOLDEN, "seq_" + seq.name + ".npz"))
    worst = 0.0
    for k in blob.files:
        if k == "residuals":
            continue  # step-to-step changes, not values: FP32 rounding dominates them near steady state
        err = _relerr(got[k], blob[k])
        worst = max(worst, err)
        assert err <= TOL_SEQ, "%s %s: relative error %.3g > %.1g" % (seq.name, k, err, TOL_SEQ)
    print("  worst rel err %.3g" % worst, flush=True)


@pytest.mark.parametrize("eq,kind", [("ns", "D2Q9"), ("ns", "D3Q15"), ("ad", "D2Q5"), ("ad", "D3Q7"),
                                     ("ad", "D2Q9")])
def test_fp32_storage_round_trip(ctx, eq, kind):
    """An equilibrium initial condition (NS: stored as deviations from the rest state) comes back to
    FP32 precision; an upload keeps it; converting to FP64 is exact and back to FP32 keeps it."""
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
    for prec in (pa.F64, pa.F32):
        lat = pa.Lattice(ctx, k, *dims, prec=prec)
        if eq == "ns":
            pa.check(ctx.lib.lbm_ns_initial_condition(lat.h, a["rho"].ptr, a["ux"].ptr, a["uy"].ptr, uzp))
        else:
            pa.check(ctx.lib.lbm_ad_initial_condition(lat.h, a["tem"].ptr, a["ux"].ptr, a["uy"].ptr, uzp))
        out[prec] = lat.download()
        if prec == pa.F32:
            pa.check(ctx.lib.lbm_lattice_set_precision(lat.h, pa.F64))
            np.testing.assert_array_equal(lat.download(), out[prec])
            pa.check(ctx.lib.lbm_lattice_set_precision(lat.h, pa.F32))
            np.testing.assert_array_equal(lat.download(), out[prec])
            lat.upload(out[pa.F64])
            assert _relerr(lat.download(), out[pa.F64]) <= 1e-7
        lat.close()
    err = _relerr(out[pa.F32], out[pa.F64])
    print("init rel err %.3g" % err, flush=True)
    assert err <= 1e-7


def test_fp32_heatsink_sensitivity_within_1e4(ctx):
    """The steady sensitivity of a small heat sink with FP32 storage carrying the solve (the default
    switch floor 1e-7 lies below eps, so FP32 runs until the solve converges or its residual
    stalls), against the FP64 steady state (FP64 to a 1e-9 residual):
      - at eps 1e-6 FP32 is within 1e-4 of it, or no farther from it than FP64 stopped at the same
        1e-6 (the two stop at different points of the slow tail; measured: FP32 1.1e-4, FP64 2.4e-4);
      - at eps 1e-8 (FP32 until it stalls, FP64 after) within 2e-5 in dfds and 1e-5 in the objective."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink2d import Heatsink2D

    out = {}
    for name, prec, eps in (("f64_tight", pa.F64, 1e-9), ("f64", pa.F64, 1e-6), ("f32", pa.F32, 1e-6),
                            ("f32_tight", pa.F32, 1e-8)):
        hs = Heatsink2D(ctx, lx=48, ly=96, ad_kind=pa.D2Q5, seed=3, filter_radius=3.0, max_steps=400000, prec=prec)
        hs.eps = eps
        r = hs.evaluate()
        out[name] = (r, hs.d["dfds"].download())
        print(name, "steps", r["fwd_steps"], r["adj_steps"], "in fp32", r.get("fp32_steps"), "objective",
              r["objective"], flush=True)
        assert max(r["fwd_residual"]) < eps and max(r["adj_residual"]) < eps
        hs.close()
    rt, st = out["f64_tight"]

    def err(name):
        r, s = out[name]
        return _relerr(s, st), abs(r["objective"] - rt["objective"]) / abs(rt["objective"])

    e64, e32, e32t = err("f64"), err("f32"), err("f32_tight")
    print("vs the FP64 steady state (dfds, objective): fp64 %s fp32 %s fp32 to 1e-8 %s" % (e64, e32, e32t), flush=True)
    r32 = out["f32"][0]
    assert r32["fp32_steps"][0] >= 0.8 * r32["fwd_steps"] and r32["fp32_steps"][1] >= 0.8 * r32["adj_steps"]
    assert e32[0] <= max(1e-4, e64[0]) and e32[1] <= max(1e-4, e64[1])
    assert e32t[0] <= 2e-5 and e32t[1] <= 1e-5


def test_fp32_divergence_is_solved_again_in_fp64(ctx):
    """Heatsink2D: a reduced-precision adjoint whose residual turns non-finite (forced from the
    host: one NaN population planted after its FP32 initial condition) is solved again from its
    cold start in FP64, as the FP64 driver would solve it on the same forward fields."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink2d import Heatsink2D

    kw = dict(lx=24, ly=40, ad_kind=pa.D2Q5, seed=1, filter_radius=2.0, use_graphs=False)
    hs = Heatsink2D(ctx, prec=pa.F32, **kw)
    orig = hs.init_adjoint

    def planted(warm=False):
        orig(warm)
        if hs.pf.prec != pa.F64:
            f = hs.pf.download()
            f[2, hs.N // 3] = np.nan
            hs.pf.upload(f)

    hs.init_adjoint = planted
    r = hs.evaluate()
    print("redo", r["fp64_redo"], r["fwd_steps"], r["adj_steps"], r["fp32_steps"], flush=True)
    assert r["fp64_redo"] == [False, True] and not r["diverged"]
    assert max(r["adj_residual"]) < 1e-6 and r["fp32_steps"][1] == 0
    got = hs.d["dfds"].download()
    # the same forward fields, adjoint solved in FP64 directly
    hs.init_adjoint = orig
    hs.in_fp64(hs.adjoint)
    hs.sensitivity()
    np.testing.assert_array_equal(got, hs.d["dfds"].download())
    hs.close()
