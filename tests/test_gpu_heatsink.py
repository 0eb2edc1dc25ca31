"""The HIP heatsink driver (panslbm2_amd/heatsink2d.py: recorded programs, HIP-graph chunks,
field rebinding, boundary-only saves) against its CPU-oracle replica (tests/heatsink_oracle.py).
Populations, fields and the objective must be value-identical; the sensitivity, whose reference
arithmetic squares with pow(), within a relative 1e-12."""
import numpy as np
import pytest

import heatsink_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kind", ["D2Q5", "D2Q9"])
@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "nographs"])
def test_heatsink_driver_matches_oracle(ctx, kind, graphs):
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink2d import Heatsink2D

    ad_kind = getattr(pa, kind)
    print("heatsink 26x18", kind, "graphs" if graphs else "eager", flush=True)
    kw = dict(lx=26, ly=18, ad_kind=ad_kind, seed=3, max_steps=400, eps=1e-30)
    hs = Heatsink2D(ctx, use_graphs=graphs, **kw)
    r = hs.evaluate()
    want = heatsink_oracle.run(**kw)
    assert r["fwd_steps"] == want["fwd_steps"] == 400 and r["adj_steps"] == want["adj_steps"] == 400
    np.testing.assert_array_equal(hs.gamma, want["gamma"])
    d = hs.d
    fb, ab = hs.fwd_buf, hs.adj_buf
    got = {"rho": d["rho"], "tem": d["tem"], "ux": d["ux%d" % fb], "uy": d["uy%d" % fb], "g": d["g"], "ip": d["ip"],
           "imx": d["imx"], "imy": d["imy"], "item": d["item"], "iqx": d["iqx%d" % ab], "iqy": d["iqy%d" % ab],
           "ig": d["ig"]}
    for k, v in got.items():
        np.testing.assert_array_equal(v.download(), want[k], err_msg=k)
    np.testing.assert_array_equal(hs.pf.download(), want["adj_f"])
    np.testing.assert_array_equal(hs.pg.download(), want["adj_g"])
    assert r["objective"] == want["objective"]
    dfds, wd = d["dfds"].download(), want["dfds"]
    assert np.max(np.abs(dfds - wd)) <= 1e-12 * np.max(np.abs(wd))
    np.testing.assert_allclose(r["fwd_residual"], want["fwd_res"], rtol=1e-10)
    hs.close()


def test_set_design_in_place_and_warm_restart(ctx):
    """Heatsink2D.set_design rewrites the coefficient arrays in place: a driver whose graphs were
    captured on design A, given design B, evaluates B exactly as a fresh eager driver given B.  A
    warm-started evaluation (from the converged solution kept by the previous one) converges in
    fewer steps to the same objective, to the spread the 1e-6 residual leaves."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink2d import Heatsink2D

    print("heatsink 2d set_design + warm", flush=True)
    kw = dict(lx=40, ly=24, ad_kind=pa.D2Q5, seed=3, filter_radius=2.0)
    a = Heatsink2D(ctx, use_graphs=True, **kw)
    a.set_design(np.full(a.N, 0.6))  # solid fraction 0.4
    a.evaluate()
    a.design_update(vf=0.4, move=0.05)  # a projected-gradient step keeps the solid fraction at vf
    assert abs((1.0 - a.x).mean() - 0.4) < 1e-6
    xb = a.x.copy()
    ra = a.evaluate()  # design B, coefficients rewritten under the graphs captured on design A
    b = Heatsink2D(ctx, use_graphs=False, **kw)
    b.set_design(xb)
    rb = b.evaluate()
    assert ra["fwd_steps"] == rb["fwd_steps"] and ra["adj_steps"] == rb["adj_steps"]
    assert ra["objective"] == rb["objective"]
    np.testing.assert_array_equal(a.d["dfds"].download(), b.d["dfds"].download())
    rw = a.evaluate(warm=True)
    print("cold", ra["fwd_steps"], ra["adj_steps"], "warm", rw["fwd_steps"], rw["adj_steps"], flush=True)
    assert rw["warm"] and rw["converged"]
    assert rw["fwd_steps"] < ra["fwd_steps"] and rw["adj_steps"] < ra["adj_steps"]
    assert abs(rw["objective"] - ra["objective"]) <= 1e-3 * abs(ra["objective"])
    a.close()
    b.close()
