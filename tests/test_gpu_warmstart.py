"""The learned warm start in the 3D heat-sink design loop (Heatsink3D.attach_warm_start).

* The warm start is a starting point only: a warm-started evaluation converges to the same
  step-to-step residual as the cold start.  The residual criterion itself leaves part of the
  slowest mode undecided, in either start: on this small, stiff duct (the 40 % solid fill raises
  the inlet density to about 25) warm and cold sensitivities stopped at 1e-6 differ by 1.1e-3 of
  their largest magnitude (bound here 5e-3) and their objectives by 7e-4 (bound 2e-3).  Converged
  to 1e-8, the same design's warm and cold sensitivities agree to 1e-4 and objectives to 1e-5.
* What the trained network gains: over the same design loop, the warm-started evaluations after
  the first need fewer LBM steps than cold starts of the same designs.
The design sequence comes from the projected-gradient update (Heatsink3D.design_update
optimizer="pg"); the MMA loop on this small, stiff duct reaches a design whose adjoint diverges
(in FP64, from the cold start too) at its sixth iteration, so the MMA loop is tested over its first
five (test_mma_design_loop_lowers_the_objective).
* A warm start whose solve diverges (forced here by a non-finite prediction) is replaced by the
  cold start, which takes the cold start's step counts and gives its sensitivities.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRID = dict(lx=32, ly=16, lz=16, plate=4, max_steps=60000)


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    return pa.Context(0)


def _max_rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_warm_start_needs_fewer_steps_and_keeps_sensitivities(ctx):
    from panslbm2_amd.heatsink3d import Heatsink3D

    print("warm start loop", GRID, flush=True)
    hs = Heatsink3D(ctx, **GRID)
    hs.attach_warm_start(seed=1, steps=60)
    iters = 6
    hist = hs.optimize(iters, vf=0.4, move=0.05, optimizer="pg")
    assert all(r["converged"] for r in hist)
    assert not hist[0]["warm"] and all(r["warm"] for r in hist[1:])
    assert not any(any(r["fallback"]) for r in hist)
    warm_steps = [r["fwd_steps"] + r["adj_steps"] for r in hist]
    # the same final design from the cold start
    x_last = hs.x.copy()
    dfds_warm = hs.d["dfds"].download()
    obj_warm = hist[-1]["objective"]
    cold = hs.evaluate(warm=False, train=False)
    dfds_cold = hs.d["dfds"].download()
    cold_steps = cold["fwd_steps"] + cold["adj_steps"]
    e6 = _max_rel(dfds_warm, dfds_cold)
    print("warm steps", warm_steps, "cold steps (last design)", cold_steps, "sample loss", hs.warm.sample_loss(),
          "dfds warm/cold at 1e-6: %.3e" % e6, flush=True)
    assert cold["converged"]
    assert e6 <= 5e-3
    assert obj_warm == pytest.approx(cold["objective"], rel=2e-3)  # measured 7e-4
    assert warm_steps[-1] < cold_steps
    # every warm-started evaluation beat the cold start of the first design
    assert max(warm_steps[1:]) < warm_steps[0]
    assert np.array_equal(hs.x, x_last)
    # converged further, warm and cold agree to 1e-4
    hs.eps = 1e-8
    w8 = hs.evaluate(warm=True, train=False)
    dfds_w8 = hs.d["dfds"].download()
    c8 = hs.evaluate(warm=False, train=False)
    dfds_c8 = hs.d["dfds"].download()
    e8 = _max_rel(dfds_w8, dfds_c8)
    print("at 1e-8: warm steps", w8["fwd_steps"] + w8["adj_steps"], "cold steps", c8["fwd_steps"] + c8["adj_steps"],
          "dfds warm/cold %.3e" % e8, flush=True)
    assert w8["converged"] and c8["converged"]
    assert e8 <= 1e-4
    assert w8["objective"] == pytest.approx(c8["objective"], rel=1e-5)
    hs.close()


def test_diverging_warm_start_falls_back_to_cold(ctx):
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, **GRID)
    ws = hs.attach_warm_start(seed=2, steps=5)
    hs.set_design(np.where(hs.k >= hs.plate, 0.6, 0.0))
    ref = hs.evaluate(warm=False, train=True)  # cold, and one observation for the network
    dfds_cold = hs.d["dfds"].download()
    assert ws.ready

    def bad_predict(inputs, outputs):
        for o in outputs:
            o.fill(np.nan)

    ws.predict = bad_predict
    r = hs.evaluate(train=False)
    assert r["warm"] and r["fallback"] == [True, True]
    assert r["converged"]
    assert r["fwd_steps"] == ref["fwd_steps"] and r["adj_steps"] == ref["adj_steps"]
    # same step counts; sensitivities agree to 5e-6 of their largest magnitude (not bit for bit)
    assert _max_rel(hs.d["dfds"].download(), dfds_cold) <= 1e-4
    assert r["objective"] == pytest.approx(ref["objective"], rel=1e-6)
    hs.close()


def test_learned_change_beats_previous_solution(ctx):
    """What the trained network gains over the method without it: the same design loop started
    from the previous converged solution alone (learn=False: the network is never trained, so its
    prediction is exactly the previous solution) needs more LBM steps than with the learned change."""
    from panslbm2_amd.heatsink3d import Heatsink3D

    steps = {}
    for learn in (False, True):
        hs = Heatsink3D(ctx, **GRID)
        hs.attach_warm_start(seed=1, steps=60, learn=learn)
        hist = hs.optimize(8, vf=0.4, move=0.05, optimizer="pg")
        assert all(r["converged"] for r in hist) and not any(any(r["fallback"]) for r in hist)
        steps[learn] = [r["fwd_steps"] + r["adj_steps"] for r in hist]
        hs.close()
    print("previous-solution steps", steps[False], "learned steps", steps[True], flush=True)
    assert steps[False][0] == steps[True][0]  # both cold
    assert sum(steps[True][2:]) < sum(steps[False][2:])


def test_mma_design_loop_lowers_the_objective(ctx):
    """The device MMA optimiser driving the 3D design loop (heatsink3D.cpp:78-85, :120-139, :282):
    every evaluation converges, the solid fraction stays at most vf, and the mean plate temperature
    falls at every iteration."""
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, **GRID)
    hist = hs.optimize(5, vf=0.4, move=0.05, warm=False, optimizer="mma")
    objs = [r["objective"] for r in hist]
    x = hs.x[hs.k >= hs.plate]
    print("MMA objectives", objs, "solid fraction", (1 - x).mean(), "interior-point iterations", hs.mma_iterations,
          flush=True)
    assert all(r["converged"] for r in hist)
    assert all(b < a for a, b in zip(objs, objs[1:]))
    assert (1 - x).mean() <= 0.4 + 1e-9
    hs.close()


def test_mma_loop_with_heaviside_projection(ctx):
    """The reference driver's loop shape (heatsink3D.cpp:103, :249-250, :282): the design is the
    Heaviside projection of the filtered variables, and MMA receives the objective and constraint
    gradients through the filter (lbm_heaviside_sensitivity).  Every evaluation converges and the
    objective falls."""
    from panslbm2_amd.heatsink3d import Heatsink3D

    hs = Heatsink3D(ctx, **GRID, heaviside_beta=2.0)
    hist = hs.optimize(4, vf=0.4, move=0.05, warm=False, optimizer="mma")
    objs = [r["objective"] for r in hist]
    print("MMA + Heaviside objectives", objs, flush=True)
    assert all(r["converged"] for r in hist)
    assert objs[-1] < objs[0]
    hs.close()
