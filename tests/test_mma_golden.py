"""The MMA fixtures (tests/golden/mma_*.npz) on the CPU: the reference harness, where it is built,
reproduces them; every step respects the move limit and the bounds (mma.h:227-231)."""
import ctypes
import os

import numpy as np
import pytest

import mmacases
from _oracle import ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("pr", mmacases.PROBLEMS, ids=lambda p: p["name"])
def test_fixture_respects_move_limits(pr):
    b = np.load(os.path.join(GOLDEN, pr["name"] + ".npz"))
    lo, hi = mmacases.xbounds(pr)
    for k in range(pr["steps"]):
        xi, xo = b["x_in/%d" % k], b["x_out/%d" % k]
        assert np.all(xo >= lo) and np.all(xo <= hi)
        assert np.abs(xo - xi).max() <= pr["move"] * (hi - lo).max() + 1e-12
        assert np.abs(xo - xi).max() > 1e-3  # the optimiser moved


@pytest.mark.skipif(ref() is None, reason="reference harness not built")
@pytest.mark.parametrize("pr", mmacases.PROBLEMS, ids=lambda p: p["name"])
def test_reference_reproduces_fixture(pr):
    R = ref()
    VP, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    R.refmma_create.restype = VP
    R.refmma_create.argtypes = [I, I, D, VP, VP, VP, VP, VP, D]
    R.refmma_update.argtypes = [VP, I, I, VP, D, VP, VP, VP]
    R.refmma_destroy.argtypes = [VP]
    p = lambda a: a.ctypes.data_as(VP)
    b = np.load(os.path.join(GOLDEN, pr["name"] + ".npz"))
    n, m = pr["n"], pr["m"]
    a, c, d = (np.array(pr[k], dtype=np.float64) for k in ("a", "c", "d"))
    lo, hi = mmacases.xbounds(pr)
    h = R.refmma_create(n, m, D(pr["a0"]), p(a), p(c), p(d), p(lo), p(hi), D(pr["move"]))
    for k in range(pr["steps"]):
        x = np.ascontiguousarray(b["x_in/%d" % k])
        R.refmma_update(h, n, m, p(x), D(float(b["f/%d" % k])), p(np.ascontiguousarray(b["dfdx/%d" % k])),
                        p(np.ascontiguousarray(b["g/%d" % k])), p(np.ascontiguousarray(b["dgdx/%d" % k])))
        np.testing.assert_array_equal(x, b["x_out/%d" % k])
    R.refmma_destroy(h)
