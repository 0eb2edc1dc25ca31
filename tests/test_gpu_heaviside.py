"""HeavisideFilter on the GPU (lbm_heaviside_filter / lbm_heaviside_sensitivity) against the
reference's fixtures (tests/golden/heaviside.npz, src/utility/heavisidefilter.h default weights).
The stencil sums run in the reference's order; tanh comes from the device's libm, so values agree
to 1e-14 relative rather than bit for bit."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("case", mg.HEAVISIDE_CASES, ids=lambda c: c[0])
def test_heaviside_matches_reference(ctx, case):
    import panslbm2_amd as pa

    name, lx, ly, lz, R, beta, seed = case
    print("heaviside case", name, flush=True)
    b = np.load(os.path.join(HERE, "golden", "heaviside.npz"))
    n = lx * ly * lz
    s, d = ctx.array((n,), data=b[name + "/s"]), ctx.array((n,), data=b[name + "/dfdrho"])
    rho, dfds = ctx.array((n,)), ctx.array((n,))
    pa.check(ctx.lib.lbm_heaviside_filter(ctx.h, lx, ly, lz, R, beta, s.ptr, rho.ptr))
    pa.check(ctx.lib.lbm_heaviside_sensitivity(ctx.h, lx, ly, lz, R, beta, s.ptr, d.ptr, dfds.ptr))
    for got, key in ((rho.download(), "rho"), (dfds.download(), "dfds")):
        want = b[name + "/" + key]
        err = np.abs(got - want).max() / np.abs(want).max()
        print("  %s rel err %.3g" % (key, err), flush=True)
        assert err <= 1e-14


@pytest.mark.parametrize("case", mg.HEAVISIDE_REGION_CASES, ids=lambda c: c[0])
def test_heaviside_region_matches_reference(ctx, case):
    """lbm_heaviside_*_region against the reference with the drivers' design-region weight lambda."""
    import panslbm2_amd as pa

    name, lx, ly, lz, R, beta, mx, my, mz, seed = case
    print("heaviside region case", name, flush=True)
    b = np.load(os.path.join(HERE, "golden", "heaviside.npz"))
    n = lx * ly * lz
    s, d = ctx.array((n,), data=b[name + "/s"]), ctx.array((n,), data=b[name + "/dfdrho"])
    rho, dfds = ctx.array((n,)), ctx.array((n,))
    pa.check(ctx.lib.lbm_heaviside_filter_region(ctx.h, lx, ly, lz, R, beta, mx, my, mz, s.ptr, rho.ptr))
    pa.check(ctx.lib.lbm_heaviside_sensitivity_region(ctx.h, lx, ly, lz, R, beta, mx, my, mz, s.ptr, d.ptr, dfds.ptr))
    for got, key in ((rho.download(), "rho"), (dfds.download(), "dfds")):
        want = b[name + "/" + key]
        err = np.abs(got - want).max() / np.abs(want).max()
        print("  %s rel err %.3g" % (key, err), flush=True)
        assert err <= 1e-14
