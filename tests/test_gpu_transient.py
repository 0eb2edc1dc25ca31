"""The transient heat-sink driver (panslbm2_amd/transient2d.py) against the reference.

tests/golden/seq_transient_nc2d_16x12.npz holds the reference's own functions composed in the
order of production/heatsink_transient.cpp:138-222 (tests/sequences.py, 16x12 grid, 60 steps):
the objective, the unfiltered heat-source sensitivity accumulated over the backward pass, and the
adjoint fields the last adjoint step leaves.  The driver runs the same problem through fused
launches, with the whole trajectory resident or with checkpoints and recomputed windows; every
value must be identical in both (FP64, the same operations in the same order).
"""
import os

import numpy as np
import pytest

import sequences

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "seq_transient_nc2d_16x12.npz")
ADJ = ("ip", "iux", "iuy", "imx", "imy", "item", "iqx", "iqy")


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _driver(ctx, **kw):
    from panslbm2_amd.transient2d import TransientHeatsink2D

    seq = [s for s in sequences.SEQUENCES if s.name == "transient_nc2d_16x12"][0]
    seq.setup()  # sets the gravity components
    N = seq.lx * seq.ly
    s = np.random.default_rng(seq.seed).uniform(0.01, 0.99, N)
    return TransientHeatsink2D(ctx, seq.lx, seq.ly, seq.nt, s=s, nu=seq.nu, pr=seq.pr, ra=seq.ra, qn=seq.qn, L=seq.L,
                               tem0=seq.tem0, gx=seq.gx, gy=seq.gy, **kw)


def _check(drv, tem_last):
    blob = np.load(GOLDEN)
    assert drv.objective == float(blob["objective"])
    np.testing.assert_array_equal(drv.d["dfds"].download(), blob["dfds"])
    for k in ADJ:
        np.testing.assert_array_equal(drv.d[k].download(), blob["adj_" + k], err_msg=k)
    np.testing.assert_array_equal(drv.d["ig"].download(), blob["adj_ig"])
    np.testing.assert_array_equal(tem_last, blob["tem_last"])


@pytest.mark.parametrize("storage,window,graphs", [("full", None, True), ("full", None, False),
                                                   ("checkpoint", None, False), ("checkpoint", 2, False),
                                                   ("checkpoint", 13, False), ("checkpoint", 59, False)])
def test_transient_adjoint_matches_reference(ctx, storage, window, graphs):
    """Twice in a row (the second evaluation replays the graphs captured by the first)."""
    print("transient", storage, window, "graphs" if graphs else "", flush=True)
    drv = _driver(ctx, storage=storage, window=window, use_graphs=graphs, chunk=16)
    assert drv.use_graphs == graphs
    for _ in range(2):
        drv.forward()
        ctx.sync()
        tem_last = drv.slots["tem"].download()[(drv.nt - 1) % (drv.nt if storage == "full" else drv.K + 1)]
        drv.adjoint()
        ctx.sync()
        _check(drv, tem_last)
    drv.close()


def test_transient_evaluate_twice_is_repeatable(ctx):
    drv = _driver(ctx, storage="checkpoint", window=5)
    f1, g1 = drv.evaluate()
    f2, g2 = drv.evaluate()
    assert f1 == f2
    np.testing.assert_array_equal(g1, g2)
    drv.close()


def test_transient_graphs_follow_a_new_design(ctx):
    """Graphs captured on design A, replayed after set_design(B): the same values as eager launches
    of design B (the coefficient arrays are rewritten in place, their addresses kept)."""
    import numpy as np

    rng = np.random.default_rng(11)
    seq = [s for s in sequences.SEQUENCES if s.name == "transient_nc2d_16x12"][0]
    N = seq.lx * seq.ly
    sb = rng.uniform(0.01, 0.99, N)
    out = {}
    for graphs in (True, False):
        drv = _driver(ctx, storage="full", use_graphs=graphs, chunk=16)
        drv.evaluate()  # design A (captures the graphs when graphs is set)
        drv.set_design(sb)
        f, dfds = drv.evaluate()
        out[graphs] = (f, np.asarray(dfds).copy())
        drv.close()
    assert out[True][0] == out[False][0]
    np.testing.assert_array_equal(out[True][1], out[False][1])
