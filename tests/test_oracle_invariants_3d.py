"""Closed-form pins for the 3D thermal boundary operators on the D3Q7 lattice, which has no reference
counterpart, on every face.

Each identity is one the reference's D3Q15 operator satisfies by construction; the D3Q15 operator
(pinned against the reference fixtures) is checked against it too, so the identity itself is
validated on the reference, and the D3Q7 operator must satisfy it as well.  Face (a, dd): axis a,
dd = -1 on a min face, +1 on a max face; U = the populations streaming left undefined
(c_a = -dd after Stream, c_a = dd after iStream), K = the opposite set.
"""
import ctypes

import numpy as np
import pytest

from _oracle import D3Q7, D3Q15, NC, Lattice, lib, ptr, to_face

D = ctypes.c_double
VEL = {
    D3Q15: ([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1], [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1],
            [0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1]),
    D3Q7: ([0, 1, 0, 0, -1, 0, 0], [0, 0, 1, 0, 0, -1, 0], [0, 0, 0, 1, 0, 0, -1]),
}
W = {D3Q15: [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8, D3Q7: [0.0] + [1 / 6] * 6}
LX, LY, LZ = 5, 4, 6
FACES = range(6)


def _c(kind):
    return np.array(VEL[kind], dtype=float)


def _setup(kind, face, seed=7):
    rng = np.random.default_rng(seed + face)
    q = Lattice(kind, LX, LY, LZ)
    N = q.N
    g = np.array(W[kind])[:, None] * rng.uniform(0.5, 1.5, (NC[kind], N))
    if kind == D3Q7:
        g[0] = rng.uniform(-0.1, 0.1, N)  # zero weight: give the rest population a value anyway
    q.set(g)
    n = np.arange(N)
    ijk = (n % LX, (n // LX) % LY, n // (LX * LY))
    a, dd = face // 2, (1 if face % 2 else -1)
    dims = (LX, LY, LZ)
    on = ijk[a] == (dims[a] - 1 if dd > 0 else 0)
    # interior of the face: off every other face, so only this face's operator acts
    for b in range(3):
        if b != a:
            on &= (ijk[b] > 0) & (ijk[b] < dims[b] - 1)
    u = [rng.uniform(-.05, .05, N) for _ in range(3)]
    return q, g, rng, on, a, dd, u


def _face_arr(kind, arr, dt=None):
    return to_face(kind, LX, LY, LZ, arr, dt)


@pytest.mark.parametrize("face", FACES)
@pytest.mark.parametrize("kind", [D3Q15, D3Q7])
def test_set_t_fixes_node_temperature(kind, face):
    """advection.h:149-238: after SetT the node's populations sum to the prescribed T."""
    q, g, rng, on, a, dd, u = _setup(kind, face)
    tem = rng.uniform(0.2, 1.0, q.N)
    lib().ol_ad_set_t(q.h, ptr(_face_arr(kind, on.astype(np.int32), np.int32)), ptr(_face_arr(kind, tem)),
                      ptr(u[0]), ptr(u[1]), ptr(u[2]))
    out = q.get()
    np.testing.assert_allclose(out.sum(0)[on], tem[on], rtol=1e-14)
    np.testing.assert_array_equal(out[:, ~on], g[:, ~on])


@pytest.mark.parametrize("face", FACES)
@pytest.mark.parametrize("kind", [D3Q15, D3Q7])
def test_set_q_flux_identity(kind, face):
    """advection.h:292-524: sum_U g = (1 - 3 dd u_a) / (1 + 3 dd u_a) ((1 + 1/(6D)) qn + sum_K g)."""
    q, g, rng, on, a, dd, u = _setup(kind, face)
    diff = rng.uniform(0.02, 0.2, q.N)
    qn = rng.uniform(-0.01, 0.01, q.N)
    lib().ol_ad_set_q(q.h, ptr(_face_arr(kind, on.astype(np.int32), np.int32)), ptr(_face_arr(kind, qn)),
                      ptr(u[0]), ptr(u[1]), ptr(u[2]), ptr(diff), D(0.0))
    out = q.get()
    ca = _c(kind)[a]
    U, K = ca == -dd, ca == dd
    ua = u[a][on]
    lhs = out[U][:, on].sum(0)
    rhs = (1 - 3 * dd * ua) / (1 + 3 * dd * ua) * ((1 + 1 / (6 * diff[on])) * qn[on] + g[K][:, on].sum(0))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-13)


def _adjoint_parts(kind, g, on, a, dd, u):
    c = _c(kind)
    w = np.array(W[kind])
    U, K = c[a] == dd, c[a] == -dd
    cu = sum(np.outer(c[b], u[b][on]) for b in range(3))
    return U, K, ((w[:, None] * (1 + 3 * cu) * g[:, on])[K]).sum(0)


@pytest.mark.parametrize("face", FACES)
@pytest.mark.parametrize("kind", [D3Q15, D3Q7])
def test_adjoint_set_t_weighted_sum_vanishes(kind, face):
    """adjointadvection.h:202-300: all unknowns equal r with sum_K w (1 + 3 c.u) g + (1 - 3 dd u_a) r / 6 = 0."""
    q, g, rng, on, a, dd, u = _setup(kind, face)
    lib().ol_aad_iset_t(q.h, ptr(_face_arr(kind, on.astype(np.int32), np.int32)), ptr(u[0]), ptr(u[1]), ptr(u[2]))
    out = q.get()
    U, K, s = _adjoint_parts(kind, g, on, a, dd, u)
    r = out[U][:, on]
    assert np.all(r == r[0])
    np.testing.assert_allclose(s + (1 - 3 * dd * u[a][on]) * r[0] / 6, 0.0, atol=1e-15)


@pytest.mark.parametrize("face", FACES)
@pytest.mark.parametrize("kind", [D3Q15, D3Q7])
def test_adjoint_set_q_identity(kind, face):
    """adjointadvection.h:368-484: sum_K w (1 + 3 c.u) g - eps/3 = (1 + 3 dd u_a) r / 6.  The reference
    drops the factor 3 of the tangential terms on ymax, zmin and zmax (faces 3-5), so there the D3Q15
    operator misses the identity by exactly those terms; D3Q7 has no tangential terms."""
    q, g, rng, on, a, dd, u = _setup(kind, face)
    eps = 0.37
    lib().ol_aad_iset_q(q.h, ptr(_face_arr(kind, on.astype(np.int32), np.int32)), ptr(u[0]), ptr(u[1]), ptr(u[2]),
                        D(eps))
    out = q.get()
    U, K, s = _adjoint_parts(kind, g, on, a, dd, u)
    r = out[U][:, on]
    assert np.all(r == r[0])
    rhs = (1 + 3 * dd * u[a][on]) * r[0] / 6
    lhs = s - eps / 3
    if kind == D3Q15 and face >= 3:
        # the unscaled tangential terms: (1 - 1/3) of sum_K w 3 c_t u_t g is missing from 12 r (1 + 3 dd u_a)
        c = _c(kind)
        w = np.array(W[kind])
        tang = sum(((w[:, None] * 3 * np.outer(c[t], u[t][on]) * g[:, on])[K]).sum(0) for t in range(3) if t != a)
        lhs = lhs - tang * (2.0 / 3.0)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-16)
