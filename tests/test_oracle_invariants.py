"""Closed-form pins for oracle functions the reference has no counterpart for (D2Q5 / D3Q7 lattices,
the D2Q5 images of the thermal boundary operators, the volumetric heat source).

Each identity below is one the reference's D2Q9 operator satisfies by construction; the D2Q9
operator (pinned against the reference fixtures) is checked against it too, so the identity itself
is validated, and the D2Q5 operator must satisfy the same identity.  A dropped term, a wrong sign
or a wrong weight in a D2Q5 formula fails these tests.
"""
import ctypes

import numpy as np
import pytest

from _oracle import D2Q5, D2Q9, D3Q7, D3Q15, NC, ND, Lattice, lib, ptr, to_face

D = ctypes.c_double
VEL = {
    D2Q9: ([0, 1, 0, -1, 0, 1, -1, -1, 1], [0, 0, 1, 0, -1, 1, 1, -1, -1], [0] * 9),
    D2Q5: ([0, 1, 0, -1, 0], [0, 0, 1, 0, -1], [0] * 5),
    D3Q15: ([0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1], [0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1],
            [0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1]),
    D3Q7: ([0, 1, 0, 0, -1, 0, 0], [0, 0, 1, 0, 0, -1, 0], [0, 0, 0, 1, 0, 0, -1]),
}
W = {D2Q9: [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, D2Q5: [1 / 3] + [1 / 6] * 4,
     D3Q15: [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8, D3Q7: [0.0] + [1 / 6] * 6}


def _c(kind):
    return np.array(VEL[kind], dtype=float)  # (3, nc)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5, D3Q15, D3Q7])
def test_weights_and_isotropy(kind):
    w = np.array(W[kind])
    c = _c(kind)
    nd = ND[kind]
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(c @ w, 0.0, atol=1e-16)
    # second moment: sum w c_a c_b = delta_ab / 3 (c_s^2 = 1/3 keeps the reference's AD constants)
    M = (c[:nd] * w) @ c[:nd].T
    np.testing.assert_allclose(M, np.eye(nd) / 3.0, atol=1e-15)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5, D3Q15, D3Q7])
def test_ad_equilibrium_moments(kind):
    """advection.h:55-69 form g_eq = w T (1 + 3 c.u): zeroth moment T, first moment T u."""
    rng = np.random.default_rng(1)
    lx, ly, lz = 5, 4, (3 if ND[kind] == 3 else 1)
    q = Lattice(kind, lx, ly, lz)
    N = q.N
    tem, ux, uy, uz = rng.uniform(0.5, 2, N), rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N)
    lib().ol_ad_init(q.h, ptr(tem), ptr(ux), ptr(uy), ptr(uz) if ND[kind] == 3 else None)
    g = q.get()
    c = _c(kind)
    np.testing.assert_allclose(g.sum(0), tem, rtol=1e-14)
    np.testing.assert_allclose(c[0] @ g, tem * ux, atol=1e-15)
    np.testing.assert_allclose(c[1] @ g, tem * uy, atol=1e-15)
    if ND[kind] == 3:
        np.testing.assert_allclose(c[2] @ g, tem * uz, atol=1e-15)


def _face_node_setup(kind, lx=6, ly=5):
    rng = np.random.default_rng(7)
    q = Lattice(kind, lx, ly)
    N = q.N
    g = np.array(W[kind])[:, None] * rng.uniform(0.5, 1.5, (NC[kind], N))
    q.set(g)
    return q, g, rng


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_set_t_fixes_node_temperature(kind):
    """advection.h:99-145: after SetT the node's populations sum to the prescribed T."""
    q, g, rng = _face_node_setup(kind)
    N, lx, ly = q.N, q.lx, q.ly
    ux, uy = rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N)
    tem = rng.uniform(0.2, 1.0, N)
    j = np.arange(N) // lx
    i = np.arange(N) % lx
    mask = (i == 0) & (j > 0) & (j < ly - 1)
    lib().ol_ad_set_t(q.h, ptr(to_face(kind, lx, ly, 1, mask.astype(np.int32), np.int32)),
                      ptr(to_face(kind, lx, ly, 1, tem)), ptr(ux), ptr(uy), None)
    out = q.get()
    np.testing.assert_allclose(out.sum(0)[mask], tem[mask], rtol=1e-14)
    np.testing.assert_array_equal(out[:, ~mask], g[:, ~mask])


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_set_q_flux_identity(kind):
    """advection.h:242-288 at x = 0: unknown populations (c_x = +1) satisfy
    sum_U g = (1 + 3 ux) / (1 - 3 ux) * ((1 + 1/(6 D)) qn + sum_K g), K the outgoing set (c_x = -1)."""
    q, g, rng = _face_node_setup(kind)
    N, lx, ly = q.N, q.lx, q.ly
    ux, uy = rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N)
    diff = rng.uniform(0.02, 0.2, N)
    qn = rng.uniform(-0.01, 0.01, N)
    i, j = np.arange(N) % lx, np.arange(N) // lx
    mask = (i == 0) & (j > 0) & (j < ly - 1)
    lib().ol_ad_set_q(q.h, ptr(to_face(kind, lx, ly, 1, mask.astype(np.int32), np.int32)),
                      ptr(to_face(kind, lx, ly, 1, qn)), ptr(ux), ptr(uy), None, ptr(diff), D(0.0))
    out = q.get()
    cx = _c(kind)[0]
    U, K = cx == 1, cx == -1
    lhs = out[U][:, mask].sum(0)
    rhs = (1 + 3 * ux[mask]) / (1 - 3 * ux[mask]) * ((1 + 1 / (6 * diff[mask])) * qn[mask] + g[K][:, mask].sum(0))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-13)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_adjoint_set_t_weighted_sum_vanishes(kind):
    """adjointadvection.h:154-198 at x = 0: all unknowns (c_x = -1) equal r with
    sum_K w (1 + 3 c.u) g + W_U (1 + 3 ux) r = 0, W_U = 1/6 the unknown weight."""
    q, g, rng = _face_node_setup(kind)
    N, lx, ly = q.N, q.lx, q.ly
    ux, uy = rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N)
    i, j = np.arange(N) % lx, np.arange(N) // lx
    mask = (i == 0) & (j > 0) & (j < ly - 1)
    lib().ol_aad_iset_t(q.h, ptr(to_face(kind, lx, ly, 1, mask.astype(np.int32), np.int32)), ptr(ux), ptr(uy), None)
    out = q.get()
    c = _c(kind)
    w = np.array(W[kind])
    U, K = c[0] == -1, c[0] == 1
    r = out[U][:, mask]
    assert np.all(r == r[0])
    cu = np.outer(c[0], ux[mask]) + np.outer(c[1], uy[mask])
    s = ((w[:, None] * (1 + 3 * cu) * g[:, mask])[K]).sum(0) + (1 / 6) * (1 + 3 * ux[mask]) * r[0]
    np.testing.assert_allclose(s, 0.0, atol=1e-15)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_adjoint_set_q_identity(kind):
    """adjointadvection.h:304-364 at x = 0: sum_K w (1 + 3 c.u) g - eps/3 = (1 - 3 ux) r / 6."""
    q, g, rng = _face_node_setup(kind)
    N, lx, ly = q.N, q.lx, q.ly
    ux, uy = rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N)
    eps = 0.37
    i, j = np.arange(N) % lx, np.arange(N) // lx
    mask = (i == 0) & (j > 0) & (j < ly - 1)
    lib().ol_aad_iset_q(q.h, ptr(to_face(kind, lx, ly, 1, mask.astype(np.int32), np.int32)), ptr(ux), ptr(uy), None,
                        D(eps))
    out = q.get()
    c = _c(kind)
    w = np.array(W[kind])
    U, K = c[0] == -1, c[0] == 1
    r = out[U][:, mask][0]
    cu = np.outer(c[0], ux[mask]) + np.outer(c[1], uy[mask])
    lhs = ((w[:, None] * (1 + 3 * cu) * g[:, mask])[K]).sum(0) - eps / 3
    np.testing.assert_allclose(lhs, (1 - 3 * ux[mask]) * r / 6, rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_adjoint_rho_flux_term(kind):
    """adjointadvection.h:488-529 (SetQ branch) at x = 0: the term added from the AD lattice is
    -T (12 sum_K w g + 18 uy sum_K w c_y g) / ((1 - 3 ux) rho), linear in g."""
    rng = np.random.default_rng(3)
    lx, ly = 6, 5
    p, q = Lattice(D2Q9, lx, ly), Lattice(kind, lx, ly)
    N = p.N
    f = np.array(W[D2Q9])[:, None] * rng.uniform(0.5, 1.5, (9, N))
    g = np.array(W[kind])[:, None] * rng.uniform(0.5, 1.5, (NC[kind], N))
    rho, ux, uy, tem = rng.uniform(.9, 1.1, N), rng.uniform(-.05, .05, N), rng.uniform(-.05, .05, N), rng.uniform(.1, 1, N)
    i, j = np.arange(N) % lx, np.arange(N) // lx
    mask = (i == 0) & (j > 0) & (j < ly - 1)
    typ = to_face(D2Q9, lx, ly, 1, (mask * 2).astype(np.int32), np.int32)
    outs = []
    for gg in (g, np.zeros_like(g)):
        p.set(f)
        q.set(gg)
        lib().ol_aad_iset_rho(p.h, q.h, ptr(typ), ptr(rho), ptr(ux), ptr(uy), None, ptr(tem), D(0.0))
        outs.append(p.get())
    flux = (outs[0] - outs[1])[3][mask]  # population 3 = (-1, 0), an unknown at x = 0
    c = _c(kind)
    w = np.array(W[kind])
    K = c[0] == 1
    s0 = (w[K, None] * g[K][:, mask]).sum(0)
    s1 = (w[K, None] * c[1][K, None] * g[K][:, mask]).sum(0)
    want = -tem[mask] * (12 * s0 + 18 * uy[mask] * s1) / ((1 - 3 * ux[mask]) * rho[mask])
    np.testing.assert_allclose(flux, want, rtol=1e-12)


@pytest.mark.parametrize("kind", [D2Q9, D2Q5])
def test_heat_source_adds_q_to_temperature(kind):
    """The volumetric source adds w_c Q to g after the Brinkman force; sum w = 1, so the saved
    temperature grows by exactly Q and the flow is untouched."""
    rng = np.random.default_rng(5)
    lx, ly = 5, 4
    N = lx * ly
    res = []
    for heat in (None, rng.uniform(0, 1e-2, N)):
        p, q = Lattice(D2Q9, lx, ly), Lattice(kind, lx, ly)
        p.set(np.array(W[D2Q9])[:, None] * rng.uniform(.9, 1.1, (9, N)) if not res else res[0]["f0"])
        q.set(np.array(W[kind])[:, None] * rng.uniform(.9, 1.1, (NC[kind], N)) if not res else res[0]["g0"])
        f0, g0 = p.get(), q.get()
        o = {k: np.zeros(N) for k in ("rho", "ux", "uy", "tem", "qx", "qy")}
        alpha, diff = np.full(N, 0.3), np.full(N, 0.05)
        lib().ol_ad_collide_fc(p.h, ptr(o["rho"]), ptr(o["ux"]), ptr(o["uy"]), None, ptr(alpha), D(0.1), q.h,
                               ptr(o["tem"]), ptr(o["qx"]), ptr(o["qy"]), None, ptr(diff), 1, None,
                               ptr(heat) if heat is not None else None)
        o.update(f0=f0, g0=g0, f=p.get(), heat=heat)
        res.append(o)
    np.testing.assert_allclose(res[1]["tem"] - res[0]["tem"], res[1]["heat"], rtol=1e-12, atol=1e-17)
    np.testing.assert_array_equal(res[1]["f"], res[0]["f"])
    np.testing.assert_array_equal(res[1]["ux"], res[0]["ux"])
