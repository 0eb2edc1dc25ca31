"""Batched launches (lbm_batch_*): the fused steps of several independent heat-sink instances in
one kernel launch give exactly what each instance's own launches give (FP64: value-identical), the
batched residual matches the single one, the autotuner's timed rounds are ordinary steps, and a
whole batched design evaluation agrees with evaluating each design alone."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [dict(lx=24, ly=40, seed=1, heat=1e-3, umax=0.05), dict(lx=32, ly=48, seed=2, heat=1.5e-3, umax=0.04),
         dict(lx=20, ly=36, seed=3, heat=5e-4, umax=0.06)]


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _instances(ctx, prec, eps=1e-6):
    from panslbm2_amd import D2Q5
    from panslbm2_amd.heatsink2d import Heatsink2D

    return [Heatsink2D(ctx, lx=s["lx"], ly=s["ly"], ad_kind=D2Q5, seed=s["seed"], heat=s["heat"], umax=s["umax"],
                       use_graphs=False, prec=prec, filter_radius=2.0, eps=eps) for s in SPECS]


def _state(h):
    d = h.d
    return [h.pf.download(), h.pg.download()] + [d[k].download() for k in ("rho", "tem", "ux0", "ux1", "qy0", "qy1")]


@pytest.mark.parametrize("prec_name,tiles", [("F64", (256, 1, 0)), ("F32", (256, 1, 0)), ("F64", (128, 4, 1)),
                                             ("F64", (64, 2, 0)), ("F32", (256, 8, 1))],
                         ids=["f64", "f32", "f64-128x4-xcd", "f64-64x2", "f32-256x8-xcd"])
def test_batched_steps_equal_single_steps(ctx, prec_name, tiles):
    """Batched steps (workgroup size, tile rows, XCD remap as given) against each instance's own
    launches: value-identical states across the check steps 99 and 100."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    prec = getattr(pa, prec_name)
    print("batched vs single steps", prec_name, tiles, flush=True)
    single = _instances(ctx, prec)
    hb = HeatsinkBatch2D(ctx, SPECS, prec=prec, use_graphs=False, tune=False, filter_radius=2.0)
    hb.adjoint = False
    for h in single + hb.hs:
        h.init_forward()
        h._fwd_step(1)
    batches = hb._make_batches(lambda h, t: h._fwd_step(t), 2)
    for b in batches.values():
        pa.check(ctx.lib.lbm_batch_configure_tiles(b, *tiles))
    for t in range(2, 106):  # crosses the check steps 99 and 100
        hb._launch(batches, t)
        for h in single:
            h._fwd_step(t)
    hb._destroy(batches)
    for a, b in zip(single, hb.hs):
        for x, y in zip(_state(a), _state(b)):
            np.testing.assert_array_equal(x, y)
    for h in single:
        h.close()
    hb.close()


def test_residual_batch_matches_single(ctx):
    import panslbm2_amd as pa

    rng = np.random.default_rng(5)
    n = 4
    lens = [100, 257, 1000, 33]
    arrs = [[ctx.array((m,), data=rng.standard_normal(m)) for _ in range(4)] for m in lens]
    a = (ctypes.c_void_p * (3 * n))()
    b = (ctypes.c_void_p * (3 * n))()
    for i, x in enumerate(arrs):
        a[3 * i], a[3 * i + 1], b[3 * i], b[3 * i + 1] = x[0].ptr, x[1].ptr, x[2].ptr, x[3].ptr
    out = np.zeros(n)
    pa.check(ctx.lib.lbm_residual_batch(ctx.h, n, 2, a, b, (ctypes.c_int * n)(*lens),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    for i, x in enumerate(arrs):
        r = ctypes.c_double()
        pa.check(ctx.lib.lbm_residual(ctx.h, 2, x[0].ptr, x[1].ptr, None, x[2].ptr, x[3].ptr, None, lens[i],
                                      ctypes.byref(r)))
        np.testing.assert_allclose(out[i], r.value, rtol=1e-12)


def test_residual_of_overflowing_field_is_nan(ctx):
    """A field whose squares overflow while its step-to-step differences do not (a diverging
    solution near 1e154): the reference's quotient is finite / inf = 0, which a convergence test
    reads as converged; lbm_residual and the batched forms return NaN.  Ordinary fields keep the
    reference's value (the oracle's restatement of residual.h)."""
    import _oracle

    import panslbm2_amd as pa

    m = 1000
    big = np.full(m, 1e154)  # each square finite, their sum over 2,000 entries not
    a0, a1 = ctx.array((m,), data=big), ctx.array((m,), data=big)
    b0, b1 = ctx.array((m,), data=big * (1 - 7e-3)), ctx.array((m,), data=big * (1 + 7e-3))
    with np.errstate(over="ignore"):
        want0 = np.sqrt(np.sum((big * 7e-3) ** 2 * 2) / np.sum(big ** 2 * 2))  # inf in the denominator: 0
    assert want0 == 0.0
    r = ctypes.c_double()
    pa.check(ctx.lib.lbm_residual(ctx.h, 2, a0.ptr, a1.ptr, None, b0.ptr, b1.ptr, None, m, ctypes.byref(r)))
    assert np.isnan(r.value)
    lens = (ctypes.c_int * 1)(m)
    a = (ctypes.c_void_p * 3)(a0.ptr, a1.ptr, None)
    b = (ctypes.c_void_p * 3)(b0.ptr, b1.ptr, None)
    out = np.zeros(1)
    pa.check(ctx.lib.lbm_residual_batch(ctx.h, 1, 2, a, b, lens, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    assert np.isnan(out[0])
    pa.check(ctx.lib.lbm_residual_batch_static(ctx.h, 1, 2, a, b, lens,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    assert np.isnan(out[0])
    rng = np.random.default_rng(9)
    x = [rng.standard_normal(m) for _ in range(4)]
    dx = [ctx.array((m,), data=v) for v in x]
    pa.check(ctx.lib.lbm_residual(ctx.h, 2, dx[0].ptr, dx[1].ptr, None, dx[2].ptr, dx[3].ptr, None, m,
                                  ctypes.byref(r)))
    want = _oracle.lib().ol_residual(2, *[v.ctypes.data_as(ctypes.c_void_p) for v in x[:2]], None,
                                     *[v.ctypes.data_as(ctypes.c_void_p) for v in x[2:]], None, m)
    assert abs(r.value - want) <= 1e-12 * want


def test_tile_tuner_rounds_are_steps(ctx):
    """lbm_batch_tune_tiles: every configuration it times is a valid one, the rounds it reports are
    the steps it ran (the instances advance exactly as many single steps), and it leaves the
    fastest configuration set."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("tile tuner", flush=True)
    single = _instances(ctx, pa.F64)
    hb = HeatsinkBatch2D(ctx, SPECS, prec=pa.F64, use_graphs=False, tune=False, filter_radius=2.0)
    hb.adjoint = False
    for h in single + hb.hs:
        h.init_forward()
        h._fwd_step(1)
    batches = hb._make_batches(lambda h, t: h._fwd_step(t), 2)
    mode = hb.hs[0]._save_mode(2, False)
    arr = (ctypes.c_void_p * 2)(batches[(0, mode)], batches[(1, mode)])
    blk, rows, xcd, rounds = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    pa.check(ctx.lib.lbm_batch_tune_tiles(arr, 2, 1, ctypes.byref(blk), ctypes.byref(rows), ctypes.byref(xcd),
                                          ctypes.byref(rounds)))
    assert blk.value in (64, 128, 256) and rows.value in (1, 2, 4, 8) and blk.value // rows.value >= 16
    assert xcd.value in (0, 1) and rounds.value == 22 * 2
    for t in range(2, 2 + 2 * rounds.value):
        for h in single:
            h._fwd_step(t)
    hb._destroy(batches)
    for a, b in zip(single, hb.hs):
        for x, y in zip(_state(a), _state(b)):
            np.testing.assert_array_equal(x, y)
    for h in single:
        h.close()
    hb.close()


def test_tuner_rounds_are_steps(ctx):
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("tuner", flush=True)
    single = _instances(ctx, pa.F64)
    hb = HeatsinkBatch2D(ctx, SPECS, prec=pa.F64, use_graphs=False, tune=False, filter_radius=2.0)
    hb.adjoint = False
    for h in single + hb.hs:
        h.init_forward()
        h._fwd_step(1)
    batches = hb._make_batches(lambda h, t: h._fwd_step(t), 2)
    mode = hb.hs[0]._save_mode(2, False)  # the save mode of the ordinary (non-residual) steps
    arr = (ctypes.c_void_p * 2)(batches[(0, mode)], batches[(1, mode)])
    blk, xcd = ctypes.c_int(), ctypes.c_int()
    pa.check(ctx.lib.lbm_batch_tune(arr, 2, 1, ctypes.byref(blk), ctypes.byref(xcd)))
    assert blk.value in (64, 128, 256) and xcd.value in (0, 1)
    for t in range(2, 2 + 2 * 6 * 2):
        for h in single:
            h._fwd_step(t)
    hb._destroy(batches)
    for a, b in zip(single, hb.hs):
        for x, y in zip(_state(a), _state(b)):
            np.testing.assert_array_equal(x, y)
    for h in single:
        h.close()
    hb.close()


@pytest.mark.parametrize("prec_name", ["F64", "F32"])
def test_batched_evaluation_matches_single(ctx, prec_name):
    """Each design's objective and sensitivity from the batch against that design evaluated alone
    in FP64.  A converged instance is frozen at its convergence check, so in FP64 it stops where it
    stops alone (the batched residual sums in another order: 1e-12).  With FP32 storage the two
    runs take different paths to the steady state, and a 1e-6 step-to-step residual leaves ~1e-4
    of the slowest mode (measured: 3e-4 in the objective), so both solve to 1e-9 here, where the
    remaining difference is the precision's: objective 1e-5, dfds 1e-4 of its largest entry."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("batched evaluation", prec_name, flush=True)
    eps = 1e-6 if prec_name == "F64" else 1e-9
    single = _instances(ctx, pa.F64, eps)
    ref = []
    for h in single:
        r = h.evaluate()
        ref.append((r["objective"], h.d["dfds"].download()))
        h.close()
    hb = HeatsinkBatch2D(ctx, SPECS, prec=getattr(pa, prec_name), use_graphs=True, tune=True, filter_radius=2.0,
                         eps=eps)
    r = hb.evaluate()
    print("batch", r["fwd_steps"], r["adj_steps"], r["fwd_conv"], r["adj_conv"], r["fp32_steps"], r["tuned"],
          flush=True)
    assert min(r["fwd_conv"]) > 0 and min(r["adj_conv"]) > 0
    tol_o, tol_s = (1e-12, 1e-12) if prec_name == "F64" else (1e-5, 1e-4)
    for (obj, dfds), h, o in zip(ref, hb.hs, r["objectives"]):
        assert abs(o - obj) <= tol_o * abs(obj)
        d = h.d["dfds"].download()
        assert np.max(np.abs(d - dfds)) <= tol_s * np.max(np.abs(dfds))
    hb.close()


def test_batched_evaluation_matches_oracle(ctx):
    """The batch in FP64 (uneven grids, graphs, the autotuner's timed rounds included) against the
    CPU-oracle replica of each design (tests/heatsink_oracle.py run): every instance stops at the
    check step the oracle stops at, with the same forward temperature and objective, and dfds
    within 1e-12 (pow-based terms)."""
    import heatsink_oracle

    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("batched evaluation vs oracle", flush=True)
    hb = HeatsinkBatch2D(ctx, SPECS, prec=pa.F64, use_graphs=True, tune=True, filter_radius=2.0, eps=1e-6)
    r = hb.evaluate()
    print("batch", r["fwd_conv"], r["adj_conv"], r["tuned"], flush=True)
    for sp, h, o, cf, ca in zip(SPECS, hb.hs, r["objectives"], r["fwd_conv"], r["adj_conv"]):
        want = heatsink_oracle.run(sp["lx"], sp["ly"], pa.D2Q5, seed=sp["seed"], heat=sp["heat"], umax=sp["umax"],
                                   filter_radius=2.0, eps=1e-6)
        print(sp, "oracle", want["fwd_steps"], want["adj_steps"], flush=True)
        assert cf == want["fwd_steps"] and ca == want["adj_steps"]
        np.testing.assert_array_equal(h.d["tem"].download(), want["tem"])
        assert o == want["objective"]
        d, wd = h.d["dfds"].download(), want["dfds"]
        assert np.max(np.abs(d - wd)) <= 1e-12 * np.max(np.abs(wd))
    hb.close()


def test_fp32_divergence_is_solved_again_in_fp64(ctx):
    """An instance whose FP32 adjoint turns non-finite (forced from the host: one NaN population
    planted after its FP32 initial condition) is solved again from its cold start in FP64.
    Every instance converges, the forward (and so the objective) is untouched, and the sensitivities
    agree with the run without the NaN to the spread of where the solves stop (the batch converts
    to FP64 when every active instance is ready, so dropping one moves that step)."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("batched fp32 with a planted NaN", flush=True)
    clean = HeatsinkBatch2D(ctx, SPECS, prec=pa.F32, use_graphs=True, tune=False, filter_radius=2.0)
    rc = clean.evaluate()
    want = [h.d["dfds"].download() for h in clean.hs]
    clean.close()
    hb = HeatsinkBatch2D(ctx, SPECS, prec=pa.F32, use_graphs=True, tune=False, filter_radius=2.0)
    h1 = hb.hs[1]
    orig = h1.init_adjoint

    def planted(warm=False):
        orig(warm)
        if h1.pf.prec != pa.F64:
            f = h1.pf.download()
            f[3, h1.N // 2] = np.nan
            h1.pf.upload(f)

    h1.init_adjoint = planted
    r = hb.evaluate()
    print("redo", r["fp64_redo"], r["adj_conv"], flush=True)
    assert r["fp64_redo"] == ([], [1])
    assert min(r["fwd_conv"]) > 0 and min(r["adj_conv"]) > 0
    assert r["objectives"] == rc["objectives"]  # the objective is the forward's
    for i, h in enumerate(hb.hs):
        d = h.d["dfds"].download()
        assert np.all(np.isfinite(d))
        assert np.max(np.abs(d - want[i])) <= 2e-3 * np.max(np.abs(want[i]))
    hb.close()


def test_batched_warm_restart(ctx):
    """HeatsinkBatch2D.evaluate(warm=True) starts every instance from its kept converged solution:
    all converge again, in fewer steps, to the same objectives within the spread of the 1e-6
    residual."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink_batch import HeatsinkBatch2D

    print("batched warm restart", flush=True)
    hb = HeatsinkBatch2D(ctx, SPECS, prec=pa.F32, use_graphs=True, tune=False, filter_radius=2.0)
    r0 = hb.evaluate()
    assert min(r0["fwd_conv"]) > 0 and min(r0["adj_conv"]) > 0
    for h in hb.hs:
        h.design_update(vf=0.4, move=0.02)
    r1 = hb.evaluate(warm=True)
    print("cold", r0["fwd_steps"], r0["adj_steps"], "warm", r1["fwd_steps"], r1["adj_steps"], flush=True)
    assert min(r1["fwd_conv"]) > 0 and min(r1["adj_conv"]) > 0
    assert r1["fwd_steps"] < r0["fwd_steps"] and r1["adj_steps"] < r0["adj_steps"]
    hb.close()
