"""The 3D heat-sink driver (panslbm2_amd/heatsink3d.py): its fused steps (one kernel pass per step)
against the same boundary calls run one by one (each value-identical to the reference, see the
3D sequence tests) -- FP64, value-identical -- and its mixed-precision evaluation against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = dict(lx=14, ly=10, lz=9, plate=3)


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _state(h):
    d = h.d
    names = ("rho", "tem", "ux0", "ux1", "uz0", "qz1", "ip", "iux0", "iuz1", "imz", "item", "iqz0")
    return [h.pf.download(), h.pg.download()] + [d[k].download() for k in names]


@pytest.mark.parametrize("ns_kind", ["D3Q15", "D3Q27"])
def test_fused_steps_equal_eager_steps(ctx, ns_kind):
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    k = getattr(pa, ns_kind)
    hs = {f: Heatsink3D(ctx, **DIMS, seed=4, use_graphs=False, fused=f, max_steps=100, ns_kind=k)
          for f in (True, False)}
    for h in hs.values():
        h.init_forward()
        for t in range(1, 121):
            h._fwd_step(t)
        h.fwd_buf = 0
        h.init_adjoint()
        for t in range(1, 61):
            h._adj_step(t, 0)
    for x, y in zip(_state(hs[True]), _state(hs[False])):
        np.testing.assert_array_equal(x, y)
    for h in hs.values():
        h.close()


@pytest.mark.parametrize("ns_kind", ["D3Q15", "D3Q27"])
def test_mixed_precision_evaluation(ctx, ns_kind):
    """FP32 storage (switching to FP64 near the end) against FP64, both solved to 1e-9 so that only
    precision separates them: objective 1e-5, dfds 1e-4 of its largest entry."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    out = {}
    for prec in (pa.F64, pa.F32):
        h = Heatsink3D(ctx, **DIMS, seed=4, prec=prec, eps=1e-9, max_steps=200000, ns_kind=getattr(pa, ns_kind))
        r = h.evaluate()
        print("3d prec", prec, r["fwd_steps"], r["adj_steps"], r["fp32_steps"], r["objective"], flush=True)
        assert not r["diverged"] and max(r["fwd_residual"]) < 1e-9 and max(r["adj_residual"]) < 1e-9
        out[prec] = (r["objective"], h.d["dfds"].download())
        h.close()
    (o64, s64), (o32, s32) = out[pa.F64], out[pa.F32]
    assert abs(o32 - o64) <= 1e-5 * abs(o64)
    assert np.max(np.abs(s32 - s64)) <= 1e-4 * np.max(np.abs(s64))


@pytest.mark.parametrize("graphs,smooth,rho_in", [(True, False, None), (False, False, None), (True, True, None),
                                                  (False, "after", None), (True, False, 1.01)],
                         ids=["graphs", "nographs", "smoothed", "smoothed_after_pass", "pressure_inlet"])
def test_heatsink3d_driver_matches_oracle(ctx, graphs, smooth, rho_in):
    """The 3D heat-sink driver (fused steps, graphs, field rebinding, boundary-only saves) against its
    CPU-oracle replica (tests/heatsink_oracle.py run3d): populations, fields and the objective
    value-identical, the sensitivity (pow-based in the reference) within 1e-12 relative.  Smoothed:
    the edges and corners concurrently with the pass (second stream, inside the step graphs), and
    after it from the pass's side states ("after")."""
    import heatsink_oracle
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    print("heatsink3d 14x10x9 vs oracle", "graphs" if graphs else "eager", "smoothed %s" % smooth if smooth else "",
          flush=True)
    after = smooth == "after"
    smooth = bool(smooth)
    kw = dict(lx=14, ly=10, lz=9, plate=3, seed=4, max_steps=400, eps=1e-30, smooth_corners=smooth, inlet_rho=rho_in)
    pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, 0 if after else 1))
    try:
        hs = Heatsink3D(ctx, use_graphs=graphs, **kw)
        r = hs.evaluate()
    finally:
        pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, 1))
    want = heatsink_oracle.run3d(ad_kind=pa.D3Q7, **kw)
    assert r["fwd_steps"] == want["fwd_steps"] == 400 and r["adj_steps"] == want["adj_steps"] == 400
    np.testing.assert_array_equal(hs.gamma, want["gamma"])
    d, fb, ab = hs.d, hs.fwd_buf, hs.adj_buf
    got = {"rho": d["rho"], "tem": d["tem"], "ux": d["ux%d" % fb], "uy": d["uy%d" % fb], "uz": d["uz%d" % fb],
           "g": d["g"], "ip": d["ip"], "imx": d["imx"], "imy": d["imy"], "imz": d["imz"], "item": d["item"],
           "iqx": d["iqx%d" % ab], "iqy": d["iqy%d" % ab], "iqz": d["iqz%d" % ab], "ig": d["ig"]}
    for k, v in got.items():
        np.testing.assert_array_equal(v.download(), want[k], err_msg=k)
    np.testing.assert_array_equal(hs.pf.download(), want["adj_f"])
    np.testing.assert_array_equal(hs.pg.download(), want["adj_g"])
    assert r["objective"] == want["objective"]
    dfds, wd = d["dfds"].download(), want["dfds"]
    assert np.max(np.abs(dfds - wd)) <= 1e-12 * np.max(np.abs(wd))
    hs.close()


def test_heatsink3d_fp16_error_vs_fp64_oracle(ctx):
    """FP16 population storage (FP32 collide, FP64 from the residual floor on) on the 3D duct: objective
    and sensitivity against the FP64 oracle replica, both run to the same step counts."""
    import heatsink_oracle
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    kw = dict(lx=14, ly=10, lz=9, plate=3, seed=4, eps=1e-9, max_steps=200000)
    want = heatsink_oracle.run3d(ad_kind=pa.D3Q7, **kw)
    hs = Heatsink3D(ctx, prec=pa.F16, **kw)
    r = hs.evaluate()
    print("fp16 3d", r["fwd_steps"], r["adj_steps"], r["fp32_steps"], r["objective"], want["objective"], flush=True)
    assert r["fp32_steps"][0] > 0 and r["fp32_steps"][1] > 0
    assert abs(r["objective"] / want["objective"] - 1.0) < 1e-6
    dfds, wd = hs.d["dfds"].download(), want["dfds"]
    assert np.max(np.abs(dfds - wd)) <= 1e-4 * np.max(np.abs(wd))
    hs.close()


def test_continuation_stages(ctx):
    """The reference driver's (qf, beta) continuation: the stage advances every stage_iterations
    iterations (or once the design stops changing) and each stage's qf reaches the Brinkman
    coefficient."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, **DIMS, seed=4, max_steps=20000, heaviside_beta=1.0)
    seen = []

    def cb(r):
        a = hs.d["alpha"].download()
        s = hs.gamma
        want = hs.alpha_max / (hs.ly - 1.0) * hs.qf * (1.0 - s) / (s + hs.qf)
        np.testing.assert_allclose(a, want, rtol=1e-14)
        seen.append((r["stage"], hs.qf, hs.heaviside_beta))

    hist = hs.optimize(6, stages=Heatsink3D.REFERENCE_STAGES, stage_iterations=2, callback=cb, optimizer="pg")
    stages = [s for s, _, _ in seen]
    assert stages == sorted(stages) and stages[0] == 0 and stages[-1] >= 2
    for st, qf, beta in seen:
        assert (qf, beta) == Heatsink3D.REFERENCE_STAGES[st]
    assert len(hist) == len(seen)
    hs.close()


def test_design_loop_matches_oracle_loop(ctx):
    """The 3D MMA design loop (Heaviside projection with the design sub-box, base plate, the
    reference driver's (qf, beta) continuation advanced every iteration) against its CPU replica
    built from the reference's HeavisideFilter and MMA and the oracle evaluation
    (tests/heatsink_oracle.py loop3d).  A given gamma evaluates value-identically; the device
    projection and MMA agree with the reference's to ~1e-15 per call, so per iteration gamma is held
    to 1e-13, the objective to 1e-10, dfds to 1e-9 and the design to 1e-9."""
    import heatsink_oracle

    from panslbm2_amd.heatsink3d import Heatsink3D

    print("3d design loop vs oracle loop", flush=True)
    kw = dict(lx=14, ly=10, lz=9, plate=3, max_steps=400, eps=1e-30)
    box = (10, 8, 7)
    hs = Heatsink3D(ctx, heaviside_beta=1.0, design_box=box, **kw)
    got = []
    hs.optimize(3, vf=0.4, move=0.05, optimizer="mma", stages=Heatsink3D.REFERENCE_STAGES, stage_iterations=1,
                callback=lambda r: got.append((r["objective"], hs.x.copy(), hs.gamma.copy(), hs.d["dfds"].download())))
    hs.close()
    want = heatsink_oracle.loop3d(14, 10, 9, 3, vf=0.4, move=0.05, plate=3, box=box, filter_radius=2.0,
                                  stages=Heatsink3D.REFERENCE_STAGES, stage_iterations=1, max_steps=400, eps=1e-30)
    assert len(got) == len(want) == 3
    for it, ((o, x, s, d), (wo, wx, ws, wd)) in enumerate(zip(got, want)):
        print("iteration", it, o, wo, np.abs(x - wx).max(), np.abs(s - ws).max(), flush=True)
        assert np.abs(s - ws).max() <= 1e-13
        assert abs(o - wo) <= 1e-10 * abs(wo)
        assert np.abs(d - wd).max() <= 1e-9 * np.abs(wd).max()
        assert np.abs(x - wx).max() <= 1e-9
    assert np.abs(got[-1][1] - got[0][1]).max() > 1e-3  # the designs moved


@pytest.mark.parametrize("prec_name", ["F32", "F16"])
def test_smoothed_concurrent_equals_after_pass_reduced_precision(ctx, prec_name):
    """With FP32 / FP16 population storage, the smoothed 3D steps in the concurrent form (second
    stream, each lane rebuilding its face neighbours' states) and in the after-pass form (the states
    the pass saved) give identical populations, fields and objective: both take the same FP64 means
    of the same face states."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    print("smoothed 3D concurrent vs after the pass", prec_name, flush=True)
    prec = getattr(pa, prec_name)
    kw = dict(lx=14, ly=10, lz=9, plate=3, seed=4, max_steps=300, eps=1e-30, smooth_corners=True, prec=prec)
    out = []
    for conc in (1, 0):
        pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, conc))
        try:
            hs = Heatsink3D(ctx, **kw)
            r = hs.evaluate()
            out.append((r["objective"], hs.pf.download(), hs.pg.download(), hs.d["dfds"].download(),
                        hs.d["tem"].download()))
            hs.close()
        finally:
            pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, 1))
    (o1, *a1), (o0, *a0) = out
    assert np.isfinite(o1) and o1 == o0
    for x, y in zip(a1, a0):
        np.testing.assert_array_equal(x, y)
