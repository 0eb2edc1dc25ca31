"""The device MMA optimiser (panslbm2_amd/csrc/mma.hip) replaying the reference optimiser's
fixtures (tests/golden/mma_*.npz, written by src/utility/mma.h through oracle/_ref).

Each step starts from the reference's design of the previous step, with the same gradients and
constraint values, so the optimiser history is the reference's.  Bar: the new design agrees with
the reference's to 1e-9 (the device sums in another order and squares / cubes by multiplication
where the reference calls pow(), which moves the interior-point iterates by ~1e-14), and the
convergence test gives the same answer.
"""
import os

import numpy as np
import pytest

import mmacases

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("pr", mmacases.PROBLEMS, ids=lambda p: p["name"])
def test_mma_matches_reference(ctx, pr):
    from panslbm2_amd.optimizer import MMA

    print("mma case", pr["name"], flush=True)
    b = np.load(os.path.join(GOLDEN, pr["name"] + ".npz"))
    n, m = pr["n"], pr["m"]
    lo, hi = mmacases.xbounds(pr)
    opt = MMA(ctx, n, m, pr["a0"], pr["a"], pr["c"], pr["d"], lo, hi, move=pr["move"])
    x = ctx.array((n,))
    dfdx = ctx.array((n,))
    dgdx = ctx.array((m, n))
    for k in range(pr["steps"]):
        x.upload(b["x_in/%d" % k])
        dfdx.upload(b["dfdx/%d" % k])
        dgdx.upload(b["dgdx/%d" % k])
        f = float(b["f/%d" % k])
        assert opt.converged(f) == bool(b["converged/%d" % k])
        it = opt.update(x, f, dfdx, b["g/%d" % k], dgdx)
        got, want = x.download(), b["x_out/%d" % k]
        err = np.abs(got - want).max()
        print("  step %d: %d interior-point iterations, max |dx| vs reference %.3g" % (k, it, err), flush=True)
        assert 0 < it < 500
        assert err <= 1e-9, (k, err)
        assert np.all(got >= lo) and np.all(got <= hi)
    opt.close()
