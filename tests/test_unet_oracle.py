"""Pins of the warm-start network's CPU restatement (oracle/unet_oracle.py), no GPU needed.

The network has no reference counterpart, so its oracle is pinned by an independent library
(torch's CPU conv3d in float64), by central finite differences of the loss, and by brute-force
definitions of pooling and upsampling.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_oracle as U  # noqa: E402


def _torch_conv(x, w, b, dims, relu):
    torch = pytest.importorskip("torch")
    nx, ny, nz = dims
    co, taps, ci = w.shape
    xt = torch.from_numpy(x.reshape(nz, ny, nx, ci).transpose(3, 0, 1, 2).copy())[None]
    if taps == 27:
        wt = torch.from_numpy(w.reshape(co, 3, 3, 3, ci).transpose(0, 4, 1, 2, 3).copy())
        pad = 1
    else:
        wt = torch.from_numpy(w.reshape(co, ci, 1, 1, 1).copy())
        pad = 0
    y = torch.nn.functional.conv3d(xt, wt, torch.from_numpy(b), padding=pad)[0]
    if relu:
        y = torch.relu(y)
    return y.numpy().transpose(1, 2, 3, 0).reshape(nx * ny * nz, co)


@pytest.mark.parametrize("taps", [27, 1])
@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 5, 3), (8, 4, 2)])
def test_conv_matches_torch(dims, taps):
    rng = np.random.default_rng(1)
    V = dims[0] * dims[1] * dims[2]
    x = rng.standard_normal((V, 5))
    w = rng.standard_normal((7, taps, 5))
    b = rng.standard_normal(7)
    for relu in (False, True):
        np.testing.assert_allclose(U.conv(x, w, b, dims, relu), _torch_conv(x, w, b, dims, relu), rtol=1e-12,
                                   atol=1e-12)


def test_conv_tap_order_is_x_fastest():
    """A weight that is one at tap (dx, dy, dz) = (1, 0, 0) shifts the field by one node in x."""
    dims = (4, 3, 2)
    V = 24
    x = np.arange(V, dtype=float)[:, None]
    w = np.zeros((1, 27, 1))
    w[0, 1 * 9 + 1 * 3 + 2, 0] = 1.0
    y = U.conv(x, w, np.zeros(1), dims, False)[:, 0]
    v = np.arange(V)
    i = v % 4
    np.testing.assert_array_equal(y, np.where(i < 3, v + 1, 0))


def test_pool_and_up_definitions():
    rng = np.random.default_rng(2)
    dims = (4, 6, 2)
    nx, ny, nz = dims
    x = rng.standard_normal((nx * ny * nz, 3))
    p = U.pool(x, dims)
    u = U.up(p, dims)
    for z in range(nz):
        for y in range(ny):
            for xx in range(nx):
                v = xx + nx * (y + ny * z)
                c = xx // 2 + (nx // 2) * (y // 2 + (ny // 2) * (z // 2))
                np.testing.assert_array_equal(u[v], p[c])
    for z in range(nz // 2):
        for y in range(ny // 2):
            for xx in range(nx // 2):
                c = xx + (nx // 2) * (y + (ny // 2) * z)
                kids = [(2 * xx + a) + nx * ((2 * y + b) + ny * (2 * z + e)) for a in (0, 1) for b in (0, 1)
                        for e in (0, 1)]
                np.testing.assert_allclose(p[c], x[kids].mean(0), rtol=1e-14)
    # adjoint identities: <pool x, g> = <x, pool_bwd g>, <up c, g> = <c, up_bwd g>
    g = rng.standard_normal(p.shape)
    assert np.isclose((p * g).sum(), (x * U.pool_bwd(g, dims)).sum(), rtol=1e-12)
    gf = rng.standard_normal(x.shape)
    assert np.isclose((u * gf).sum(), (p * U.up_bwd(gf, dims)).sum(), rtol=1e-12)


def test_conv_bwd_is_adjoint():
    rng = np.random.default_rng(3)
    dims = (4, 4, 4)
    x = rng.standard_normal((64, 3))
    w = rng.standard_normal((5, 27, 3))
    g = rng.standard_normal((64, 5))
    dx, dw, db = U.conv_bwd(g, x, w, dims)
    y = U.conv(x, w, np.zeros(5), dims, False)
    # linear in x and in w: <g, conv(x)> = <dx, x> = <dw, w>
    s = (g * y).sum()
    assert np.isclose(s, (dx * x).sum(), rtol=1e-12)
    assert np.isclose(s, (dw * w).sum(), rtol=1e-12)
    np.testing.assert_allclose(db, g.sum(0), rtol=1e-14)


def test_full_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    dims = (4, 4, 4)
    net = U.UNet(dims, cin=2, cout=3, base=2)
    params = rng.standard_normal(net.nparams) * 0.5
    x0 = rng.standard_normal((64, 2))
    target = rng.standard_normal((64, 3))
    loss, g = net.grad(params, x0, target)
    assert loss == pytest.approx(net.loss(params, x0, target))
    idx = rng.choice(net.nparams, 60, replace=False)
    # include every layer's first weight and bias
    idx = np.unique(np.concatenate([idx, [l[4] for l in net.layers], [l[5] for l in net.layers]]))
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (net.loss(p1, x0, target) - net.loss(p2, x0, target)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])  # ReLU kinks aside, ~1e-7


def test_bf16_rounding():
    assert U.bf16(1.0) == 1.0
    assert U.bf16(1.0 + 2.0 ** -8) == 1.0  # tie: to even
    assert U.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert U.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    assert U.bf16(-3.0) == -3.0


def test_adam_first_step_moves_by_lr():
    """After one step with bias correction each parameter moves by lr sign(g) (for |g| >> eps)."""
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    m, v = np.zeros(3), np.zeros(3)
    U.adam(p, g, m, v, 1, lr=0.01)
    np.testing.assert_allclose(p, [0.99, -1.99, 0.49], rtol=1e-6)
