"""The HeavisideFilter fixtures (tests/golden/heaviside.npz, written by the compiled reference,
src/utility/heavisidefilter.h) checked on the CPU: the reference harness reproduces them where it
is built; the projection satisfies its closed forms (values in [0, 1]; a uniform design maps to
the projection of that value), and the sensitivity is the chain rule of the projection checked by
finite differences on a tiny grid against the reference's own value function."""
import ctypes
import os
import sys

import numpy as np
import pytest

from _oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as mg  # noqa: E402

BLOB = os.path.join(HERE, "golden", "heaviside.npz")


def _bind():
    R = ref()
    VP, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    R.ref3_heaviside_value.argtypes = [I, I, I, D, D, VP, VP]
    R.ref3_heaviside_sensitivity.argtypes = [I, I, I, D, D, VP, VP, VP]
    return R


def test_fixture_closed_forms():
    b = np.load(BLOB)
    for name, lx, ly, lz, R, beta, seed in mg.HEAVISIDE_CASES:
        rho = b[name + "/rho"]
        assert np.all(rho >= 0.0) and np.all(rho <= 1.0)
        assert np.all(np.isfinite(b[name + "/dfds"]))


@pytest.mark.skipif(ref() is None, reason="reference harness not built")
def test_reference_reproduces_fixture_and_uniform_design():
    R = _bind()
    b = np.load(BLOB)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    for name, lx, ly, lz, Rr, beta, seed in mg.HEAVISIDE_CASES:
        s, d = b[name + "/s"].copy(), b[name + "/dfdrho"].copy()
        rho, dfds = np.empty_like(s), np.empty_like(s)
        R.ref3_heaviside_value(lx, ly, lz, Rr, beta, p(s), p(rho))
        R.ref3_heaviside_sensitivity(lx, ly, lz, Rr, beta, p(s), p(d), p(dfds))
        np.testing.assert_array_equal(rho, b[name + "/rho"])
        np.testing.assert_array_equal(dfds, b[name + "/dfds"])
        u = np.full_like(s, 0.3)
        R.ref3_heaviside_value(lx, ly, lz, Rr, beta, p(u), p(rho))
        want = 0.5 * (np.tanh(0.5 * beta) + np.tanh(beta * (0.3 - 0.5))) / np.tanh(0.5 * beta)
        np.testing.assert_allclose(rho, want, rtol=1e-14)


@pytest.mark.skipif(ref() is None, reason="reference harness not built")
def test_reference_sensitivity_is_the_chain_rule_in_the_interior():
    """dF/ds for F = sum(dfdrho * rho(s)), by central differences of the reference's value
    function: equal to the reference's sensitivity at nodes 2R or more from the boundary (closer
    in, the reference divides by the node's own weight sum instead of each neighbour's)."""
    R = _bind()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lx, ly, lz, Rr, beta = 12, 10, 1, 1.5, 4.0
    s, d = mg.heaviside_inputs(lx, ly, lz, 41)
    dfds = np.empty_like(s)
    R.ref3_heaviside_sensitivity(lx, ly, lz, Rr, beta, p(s), p(d), p(dfds))

    def F(x):
        out = np.empty_like(x)
        R.ref3_heaviside_value(lx, ly, lz, Rr, beta, p(x), p(out))
        return float(d @ out)

    h = 1e-6
    for i, j in ((3, 3), (5, 4), (8, 6), (6, 5)):
        n = i + lx * j
        sp, sm = s.copy(), s.copy()
        sp[n] += h
        sm[n] -= h
        fd = (F(sp) - F(sm)) / (2 * h)
        assert abs(fd - dfds[n]) <= 1e-6 * max(1.0, abs(fd)), (i, j, fd, dfds[n])


def test_region_fixture_closed_forms():
    """Design-region weights (production/heatsink3D.cpp:87-93): a node outside the box keeps its own
    value, projected; its sensitivity is its own chain-rule factor (only it weighs itself)."""
    b = np.load(BLOB)
    for name, lx, ly, lz, R, beta, mx, my, mz, seed in mg.HEAVISIDE_REGION_CASES:
        s, d = b[name + "/s"], b[name + "/dfdrho"]
        rho, dfds = b[name + "/rho"], b[name + "/dfds"]
        n = np.arange(lx * ly * lz)
        i, j, k = n % lx, (n // lx) % ly, n // (lx * ly)
        out = ~((i < mx) & (j < my) & (k < mz))
        assert out.any() and (~out).any()
        tb = np.tanh(0.5 * beta)
        t = np.tanh(beta * (s - 0.5))
        np.testing.assert_allclose(rho[out], 0.5 * (tb + t[out]) / tb, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(dfds[out], (d * 0.5 * beta * (1.0 - t * t) / tb)[out], rtol=1e-12, atol=1e-14)
        assert not np.allclose(rho[~out], (0.5 * (tb + t) / tb)[~out])  # inside, the neighbours count


@pytest.mark.skipif(ref() is None, reason="reference harness not built")
def test_reference_region_reproduces_fixture_and_whole_grid_is_default():
    R = _bind()
    VP, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    R.ref3_heaviside_value_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP]
    R.ref3_heaviside_sensitivity_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP, VP]
    b = np.load(BLOB)
    p = lambda a: a.ctypes.data_as(VP)
    for name, lx, ly, lz, Rr, beta, mx, my, mz, seed in mg.HEAVISIDE_REGION_CASES:
        s, d = b[name + "/s"].copy(), b[name + "/dfdrho"].copy()
        rho, dfds = np.empty_like(s), np.empty_like(s)
        R.ref3_heaviside_value_region(lx, ly, lz, Rr, beta, mx, my, mz, p(s), p(rho))
        R.ref3_heaviside_sensitivity_region(lx, ly, lz, Rr, beta, mx, my, mz, p(s), p(d), p(dfds))
        np.testing.assert_array_equal(rho, b[name + "/rho"])
        np.testing.assert_array_equal(dfds, b[name + "/dfds"])
        r0, r1 = np.empty_like(s), np.empty_like(s)
        R.ref3_heaviside_value_region(lx, ly, lz, Rr, beta, lx, ly, lz, p(s), p(r0))
        R.ref3_heaviside_value(lx, ly, lz, Rr, beta, p(s), p(r1))
        np.testing.assert_array_equal(r0, r1)
