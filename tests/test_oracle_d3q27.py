"""Pins of the D3Q27 lattice and its boundary operators (no reference counterpart), oracle only.

The D3Q27 velocity and pressure boundaries are the non-equilibrium bounce-back construction the
reference's D3Q15 Zou-He operators are an instance of (oracle ol_nebb_3d).  Pinned by:
  * the construction run on D3Q15 equals the reference-pinned D3Q15 operators, all four
    (velocity, pressure, and their adjoints), on every face;
  * on D3Q27: the density and momentum the forward operators must produce (all three
    components), and equilibrium populations are a fixed point;
  * the adjoint operators equal the transpose of the forward operator's Jacobian on the outgoing
    populations (plus the pressure objective's source), computed from the forward operator by
    linearity; the same check passes for the reference's own D3Q15 pair;
  * the lattice itself: equilibrium moments up to second order, and the 4th-order isotropy.
"""
import ctypes

import numpy as np
import pytest

from _oracle import D3Q15, D3Q27, NC, Lattice, lib, ptr, to_face

D = ctypes.c_double
LX, LY, LZ = 5, 4, 6
FACES = range(6)
C27 = (
    [0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1],
    [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, 1, -1, -1, 1, -1],
    [0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1, 1, 1, 1, -1, -1, -1, -1, 1],
)
C15 = ([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1], [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1],
       [0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1])
VEL = {D3Q15: np.array(C15, dtype=float), D3Q27: np.array(C27, dtype=float)}
W = {D3Q15: np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8),
     D3Q27: np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)}


def _face(kind, a, dt=None):
    return to_face(kind, LX, LY, LZ, a, dt)


def _geometry(face):
    n = np.arange(LX * LY * LZ)
    ijk = (n % LX, (n // LX) % LY, n // (LX * LY))
    a, dd = face // 2, (1 if face % 2 else -1)
    dims = (LX, LY, LZ)
    on = ijk[a] == (dims[a] - 1 if dd > 0 else 0)
    for b in range(3):
        if b != a:
            on &= (ijk[b] > 0) & (ijk[b] < dims[b] - 1)
    return on, a, dd


def _pops(kind, seed):
    rng = np.random.default_rng(seed)
    return W[kind][:, None] * rng.uniform(0.5, 1.5, (NC[kind], LX * LY * LZ))


def _run(kind, g, on, op, v, eps=0.0, generic=True):
    """Apply operator op (0 SetU, 1 SetRho, 2 iSetU, 3 iSetRho) on the masked face nodes."""
    q = Lattice(kind, LX, LY, LZ)
    q.set(g)
    mask = _face(kind, on.astype(np.int32), np.int32)
    fv = [_face(kind, np.ascontiguousarray(x, dtype=np.float64)) for x in v]
    L = lib()
    if generic:
        L.ol_nebb_3d(q.h, op, ptr(mask), ptr(fv[0]), ptr(fv[1]), ptr(fv[2]), D(eps))
    elif op == 0:
        L.ol_ns_set_u(q.h, ptr(mask), ptr(fv[0]), ptr(fv[1]), ptr(fv[2]))
    elif op == 1:
        L.ol_ns_set_rho(q.h, ptr(mask), ptr(fv[0]), ptr(fv[1]), ptr(fv[2]))
    elif op == 2:
        L.ol_ans_iset_u(q.h, ptr(mask), ptr(fv[0]), ptr(fv[1]), ptr(fv[2]), D(eps), 1)
    else:
        L.ol_ans_iset_rho(q.h, ptr(mask))
    return q.get()


def _inputs(op, seed):
    rng = np.random.default_rng(seed)
    N = LX * LY * LZ
    if op == 1:  # rho, and the two tangential velocities
        return [rng.uniform(0.95, 1.05, N), rng.uniform(-0.05, 0.05, N), rng.uniform(-0.05, 0.05, N)]
    return [rng.uniform(-0.05, 0.05, N) for _ in range(3)]


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("face", FACES)
def test_construction_is_the_reference_on_d3q15(face, op):
    on, _, _ = _geometry(face)
    g = _pops(D3Q15, 10 + face)
    v = _inputs(op, 20 + face)
    got = _run(D3Q15, g, on, op, v, eps=0.37, generic=True)
    want = _run(D3Q15, g, on, op, v, eps=0.37, generic=False)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-16)
    assert not np.array_equal(want, g)  # the operator acted


@pytest.mark.parametrize("face", FACES)
def test_d3q27_set_u_gives_density_and_momentum(face):
    on, a, dd = _geometry(face)
    c = VEL[D3Q27]
    g = _pops(D3Q27, face)
    u = _inputs(0, 30 + face)
    out = _run(D3Q27, g, on, 0, u)
    known = g[c[a] == 0].sum(0) + 2 * g[c[a] == dd].sum(0)
    rho = known / (1 + dd * u[a])
    np.testing.assert_allclose(out.sum(0)[on], rho[on], rtol=1e-13)
    for ax in range(3):
        np.testing.assert_allclose((c[ax][:, None] * out).sum(0)[on], (rho * u[ax])[on], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(out[:, ~on], g[:, ~on])
    np.testing.assert_array_equal(out[c[a] != -dd], g[c[a] != -dd])  # only the unknowns change


@pytest.mark.parametrize("face", FACES)
def test_d3q27_set_rho_gives_density_and_tangential_momentum(face):
    on, a, dd = _geometry(face)
    c = VEL[D3Q27]
    g = _pops(D3Q27, 40 + face)
    rho, us, ut = _inputs(1, 50 + face)
    out = _run(D3Q27, g, on, 1, [rho, us, ut])
    np.testing.assert_allclose(out.sum(0)[on], rho[on], rtol=1e-13)
    t1, t2 = (a + 1) % 3, (a + 2) % 3
    for t, ut_ in ((t1, us), (t2, ut)):
        np.testing.assert_allclose((c[t][:, None] * out).sum(0)[on], (rho * ut_)[on], rtol=1e-12, atol=1e-15)


def _equilibrium(kind, rho, u):
    c, w = VEL[kind], W[kind]
    cu = sum(c[ax][:, None] * u[ax][None] for ax in range(3))
    uu = sum(x * x for x in u)
    return w[:, None] * rho[None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu[None])


def test_d3q27_equilibrium_moments():
    """Through ol_ns_init: density, momentum and the second moment rho (u u + I/3) exactly."""
    rng = np.random.default_rng(3)
    N = LX * LY * LZ
    rho = rng.uniform(0.9, 1.1, N)
    u = [rng.uniform(-0.05, 0.05, N) for _ in range(3)]
    q = Lattice(D3Q27, LX, LY, LZ)
    lib().ol_ns_init(q.h, ptr(rho), ptr(u[0]), ptr(u[1]), ptr(u[2]))
    f = q.get()
    np.testing.assert_allclose(f, _equilibrium(D3Q27, rho, u), rtol=1e-13)
    c = VEL[D3Q27]
    np.testing.assert_allclose(f.sum(0), rho, rtol=1e-13)
    for a in range(3):
        np.testing.assert_allclose((c[a][:, None] * f).sum(0), rho * u[a], rtol=1e-11, atol=1e-16)
        for b in range(3):
            pi = (c[a][:, None] * c[b][:, None] * f).sum(0)
            np.testing.assert_allclose(pi, rho * (u[a] * u[b] + (a == b) / 3.0), rtol=1e-12, atol=1e-15)
    w = W[D3Q27]
    assert np.isclose((w * c[0] ** 2 * c[1] ** 2).sum(), 1 / 9) and np.isclose((w * c[0] ** 4).sum(), 1 / 3)


@pytest.mark.parametrize("face", FACES)
def test_d3q27_equilibrium_is_a_fixed_point(face):
    on, a, dd = _geometry(face)
    rng = np.random.default_rng(60 + face)
    N = LX * LY * LZ
    rho = rng.uniform(0.9, 1.1, N)
    u = [rng.uniform(-0.05, 0.05, N) for _ in range(3)]
    feq = _equilibrium(D3Q27, rho, u)
    out = _run(D3Q27, feq, on, 0, u)
    np.testing.assert_allclose(out, feq, rtol=1e-12, atol=1e-16)
    out = _run(D3Q27, feq, on, 1, [rho, u[(a + 1) % 3], u[(a + 2) % 3]])
    np.testing.assert_allclose(out, feq, rtol=1e-12, atol=1e-16)


@pytest.mark.parametrize("kind", [D3Q15, D3Q27])
@pytest.mark.parametrize("adj_op", [2, 3])
@pytest.mark.parametrize("face", FACES)
def test_adjoint_is_the_transpose_of_the_forward(kind, adj_op, face):
    """f*_k (k outgoing) = sum over the unknowns c of f*_c d f_c / d f_k, minus the pressure
    objective's source 2 eps / (3 (1 + dd u_a)) for the velocity operator.  D3Q15 uses the
    reference-pinned operators, D3Q27 the construction."""
    on, a, dd = _geometry(face)
    c = VEL[kind]
    generic = kind == D3Q27
    fwd_op = adj_op - 2
    v = _inputs(fwd_op, 70 + face)
    g0 = _pops(kind, 80 + face)
    base = _run(kind, g0, on, fwd_op, v, generic=generic)
    U = np.nonzero(c[a] == -dd)[0]
    K = np.nonzero(c[a] == dd)[0]
    jac = {}
    for k in K:  # the forward operator is affine in the populations (u or rho fixed)
        g1 = g0.copy()
        g1[k] += 1.0
        jac[k] = _run(kind, g1, on, fwd_op, v, generic=generic)[U] - base[U]
    eps = 0.37 if adj_op == 2 else 0.0
    fs = _pops(kind, 90 + face) - 0.5 * W[kind][:, None]
    out = _run(kind, fs, on, adj_op, v, eps=eps, generic=generic)
    src = 2.0 * eps / (3.0 * (1.0 + dd * v[a])) if adj_op == 2 else 0.0
    for k in K:
        want = (fs[U] * jac[k]).sum(0) - src
        np.testing.assert_allclose(out[k][on], want[on], rtol=1e-10, atol=1e-14)
    rest = np.setdiff1d(np.arange(NC[kind]), K)
    np.testing.assert_array_equal(out[rest], fs[rest])


def test_oracle_runs_every_3d_function_on_d3q27():
    """The oracle side of the D3Q27 GPU parity cases: every 3D function runs on D3Q27 + D3Q7 and
    gives finite values, and the D3Q27 case is not a copy of a D3Q15 one."""
    import refcases

    for case in refcases.CASES_D3Q27:
        inp = refcases.make_inputs(case)
        assert inp["f"].shape[0] == 27 and inp["g"].shape[0] == 7
        for op in refcases.ops_for(case):
            out = refcases.run_oracle(case, op, inp)
            for k, v in out.items():
                assert np.all(np.isfinite(v)), (case.name, op, k)
