"""Pins of the BGK-only convection steps and the mass-flow adjoint step, oracle only.

The per-call golden fixtures (tests/golden/*.npz, written by the compiled reference) pin
ol_ad_collide_bgk and ol_aad_collide_mf value for value; these checks tie them to what the
mathematics fixes:
  * the BGK forced step is the Brinkman forced step with alpha = 0 and a uniform diffusivity;
  * BGK collisions conserve mass and temperature (sum f, sum g), with or without the buoyancy
    force, whose mass moment is zero;
  * the mass-flow step with a zero direction field is the natural-convection step; its outputs are
    affine in the direction field (the source is linear in d and everything after it is linear in
    the populations), so doubling d doubles the change.
"""
import ctypes

import numpy as np
import pytest

import refcases
from _oracle import NC, ND, Lattice, lib, ptr

D = ctypes.c_double
CASES = refcases.CASES_2D + refcases.CASES_3D


def _lat(case, kind, pops):
    h = Lattice(kind, case.lx, case.ly, case.lz)
    h.set(pops)
    return h


def _bgk(case, inp, natural, diff):
    p, q = _lat(case, case.kind, inp["f"]), _lat(case, case.kind_b, inp["g"])
    o = {k: np.zeros(case.N) for k in ("rho", "ux", "uy", "uz", "tem", "qx", "qy", "qz")}
    lib().ol_ad_collide_bgk(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), D(0.1), q.h, ptr(o["tem"]),
                            ptr(o["qx"]), ptr(o["qy"]), ptr(o["qz"]), D(diff), natural,
                            *[D(x) for x in refcases.NC_G], D(refcases.NC_TEM0), 1)
    o["f"], o["g"] = p.get(), q.get()
    return o


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_bgk_forced_step_is_brinkman_step_without_alpha(case):
    inp = refcases.make_inputs(case)
    dc = float(inp["dconst"])
    got = _bgk(case, inp, 0, dc)
    mod = dict(inp, alpha=np.zeros(case.N), diff=np.full(case.N, dc))
    want = refcases.run_oracle(case, "ad_collide_fc", mod)
    for k in ("f", "g", "rho", "ux", "uy", "tem", "qx", "qy"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-14, atol=1e-16, err_msg=k)


@pytest.mark.parametrize("natural", [0, 1])
@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_bgk_steps_conserve_mass_and_temperature(case, natural):
    inp = refcases.make_inputs(case)
    got = _bgk(case, inp, natural, float(inp["dconst"]))
    np.testing.assert_allclose(got["f"].sum(0), inp["f"].sum(0), rtol=1e-14)
    np.testing.assert_allclose(got["g"].sum(0), inp["g"].sum(0), rtol=1e-14)
    if natural:  # the buoyancy force acted
        assert not np.allclose(got["f"], _bgk(case, inp, 0, float(inp["dconst"]))["f"], rtol=1e-12, atol=0)


def _mf(case, inp, dirs):
    p, q = _lat(case, case.kind, inp["f"]), _lat(case, case.kind_b, inp["g"])
    names = ("ip", "iux", "iuy", "iuz", "imx", "imy", "imz", "item", "iqx", "iqy", "iqz")
    o = {k: np.zeros(case.N) for k in names}
    snap = np.zeros((NC[case.kind_b], case.N))
    uz = inp["uz"] if ND[case.kind] == 3 else None
    lib().ol_aad_collide_mf(p.h, ptr(inp["rho"]), ptr(inp["ux"]), ptr(inp["uy"]), ptr(uz),
                            *[ptr(o[k]) for k in names[:7]], ptr(inp["alpha"]), D(0.1), q.h, ptr(inp["tem"]),
                            *[ptr(o[k]) for k in names[7:]], ptr(inp["diff"]), *[D(x) for x in refcases.NC_G],
                            ptr(dirs[0]), ptr(dirs[1]), ptr(dirs[2]) if uz is not None else None, 1, ptr(snap))
    o["f"], o["g"], o["igsnap"] = p.get(), q.get(), snap
    return o


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_mass_flow_step_reduces_and_is_affine_in_direction(case):
    inp = refcases.make_inputs(case)
    zero = [np.zeros(case.N)] * 3
    d1 = [np.ascontiguousarray(x) for x in refcases.mf_dirs(inp)]
    d2 = [2.0 * x for x in d1]
    o0, o1, o2 = _mf(case, inp, zero), _mf(case, inp, d1), _mf(case, inp, d2)
    nc = refcases.run_oracle(case, "aad_collide_nc", inp)
    for k in nc:
        np.testing.assert_array_equal(o0[k], nc[k], err_msg=k)
    for k in ("f", "g", "iux", "imx", "item"):  # not ip, iq: the sources carry no such moment
        d_1, d_2 = o1[k] - o0[k], o2[k] - o0[k]
        assert np.abs(d_1).max() > 1e-6, k  # the source acted
        np.testing.assert_allclose(d_2, 2.0 * d_1, rtol=1e-9, atol=1e-13 * np.abs(o0[k]).max(), err_msg=k)
