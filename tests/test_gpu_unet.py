"""The warm-start network on the GPU (panslbm2_amd/csrc/unet.hip) against its CPU restatement.

Tolerances.  The device stores weights, activations and the gradients it feeds the matrix cores
in bfloat16 (8 significant bits) and accumulates in FP32.  Against the oracle run with the same
bfloat16 roundings (bf=True) the only differences are FP32 vs FP64 sums and the rare rounding that
lands on the other side of a bfloat16 tie, so every tensor agrees to 2^-7 of its largest
magnitude.  Against the plain FP64 oracle (bf=False) the bfloat16 storage itself is the error.
On these small random problems the bfloat16-emulating oracle and the FP64 oracle differ by up to
0.11 in relative L2 norm per layer gradient (cosine >= 0.994, measured over four seeds and both
grids), so the FP64 bound is 0.2 with cosine > 0.98; it catches a gradient of the wrong layer or
sign, while the bfloat16 comparison pins the kernels.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_oracle as U  # noqa: E402

pytestmark = pytest.mark.gpu

CH = {"e1a": 16, "e1b": 16, "p1": 16, "e2a": 32, "e2b": 32, "p2": 32, "e3a": 64, "e3b": 64, "cat2": 96, "d2a": 32,
      "d2b": 32, "cat1": 48, "d1a": 16, "d1b": 16}


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    return pa.Context(0)


def _setup(ctx, dims, cin, cout, seed):
    from panslbm2_amd.warmstart import UNet

    print("unet case dims=%s cin=%d cout=%d seed=%d" % (dims, cin, cout, seed), flush=True)
    rng = np.random.default_rng(seed)
    V = dims[0] * dims[1] * dims[2]
    net = UNet(ctx, dims, cin, cout, seed=seed)
    ref = U.UNet(dims, cin, cout)
    assert net.nparams == ref.nparams
    p = net.params()  # the device's own He-uniform start, as FP32 values (the head starts at zero)
    # a nonzero head and nonzero biases, so that every gradient path is exercised
    name, ci, co, taps, wo, bo = ref.layers[-1]
    p[wo:bo] = np.float32(rng.uniform(-1, 1, bo - wo) * np.sqrt(3.0 / ci))
    for name, ci, co, taps, wo, bo in ref.layers:
        p[bo:bo + co] = np.float32(rng.uniform(-0.05, 0.05, co))
    net.set_params(p)
    fin = [rng.uniform(0, 1, V) for _ in range(cin)]
    fout = [rng.standard_normal(V) * 0.5 + 0.1 * c for c in range(cout)]
    din = [ctx.array((V,), data=f) for f in fin]
    dout = [ctx.array((V,), data=f) for f in fout]
    return net, ref, p, np.stack(fin, 1), np.stack(fout, 1), din, dout


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 8, 4)])
def test_forward_matches_oracle(ctx, dims):
    net, ref, p, x0, _, din, _ = _setup(ctx, dims, 2, 13, 1)
    net.set_input(din)
    net.forward()
    a = ref.forward(p, x0, bf=True)
    for name, c in CH.items():
        got, want = net.read(name, c), a[name]
        tol = 2.0 ** -7 * max(np.abs(want).max(), 1e-30)
        err = np.abs(got - want).max()
        assert err <= tol, "%s: max error %.3e, tolerance %.3e" % (name, err, tol)
    y = net.read("y", 13)
    err = np.abs(y - a["y"]).max() / np.abs(a["y"]).max()
    assert err <= 2.0 ** -7, err
    # against plain FP64: the bfloat16 storage error only
    y64 = ref.forward(p, x0, bf=False)["y"]
    assert np.linalg.norm(y - y64) / np.linalg.norm(y64) < 3e-2
    net.close()


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 8, 4)])
def test_gradients_match_oracle(ctx, dims):
    net, ref, p, x0, t, din, dout = _setup(ctx, dims, 2, 13, 2)
    net.set_input(din)
    net.set_target(dout)
    net.forward()
    loss = net.backward()
    g = net.grads()
    lb, gb = ref.grad(p, x0, t, bf=True)
    l64, g64 = ref.grad(p, x0, t, bf=False)
    assert loss == pytest.approx(lb, rel=1e-3)
    assert loss == pytest.approx(l64, rel=2e-2)
    for name, ci, co, taps, wo, bo in ref.layers:
        for lo, hi, what in ((wo, bo, "W"), (bo, bo + co, "b")):
            a, b, c = g[lo:hi], gb[lo:hi], g64[lo:hi]
            scale = max(np.abs(b).max(), 1e-30)
            assert np.abs(a - b).max() <= 2.0 ** -6 * scale, "%s.%s vs bf16 oracle: %.3e" % (
                name, what, np.abs(a - b).max() / scale)
            rel = np.linalg.norm(a - c) / max(np.linalg.norm(c), 1e-30)
            cos = a @ c / max(np.linalg.norm(a) * np.linalg.norm(c), 1e-300)
            assert rel < 0.2 and cos > 0.98, "%s.%s vs FP64 oracle: rel %.3e cos %.6f" % (name, what, rel, cos)
    net.close()


def test_adam_step_matches_oracle(ctx):
    dims = (8, 8, 4)
    net, ref, p, x0, t, din, dout = _setup(ctx, dims, 1, 5, 3)
    net.set_input(din)
    net.set_target(dout)
    net.forward()
    net.backward()
    g = net.grads()
    net.adam(1e-3)
    got = net.params()
    want = p.copy()
    U.adam(want, g, np.zeros_like(p), np.zeros_like(p), 1, lr=1e-3)
    # FP32 master weights (two FP32 spacings of each value) and FP32 hyper-parameters: 1 - 0.999f
    # is 1e-3 (1 + 4.7e-5), so the update itself carries up to about 1e-4 of lr
    tol = 2 * np.spacing(np.abs(want).astype(np.float32)).astype(np.float64) + 1e-4 * 1e-3
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()
    big = np.abs(g) > 1e-5  # a parameter with |g| >> eps moves by lr on the first step
    assert big.sum() > 1000
    np.testing.assert_allclose(np.abs(got - p)[big], 1e-3, rtol=2e-3)
    net.close()


def test_training_fits_a_field(ctx):
    """Adam on one (input, target) pair: the loss falls by 10x in 150 steps and the output then
    approximates the target; two networks with the same seed train bit for bit alike."""
    dims = (8, 8, 8)
    V = 512
    n = np.arange(V)
    x, y, z = n % 8, (n // 8) % 8, n // 64
    inp = [ctx.array((V,), data=(x + y) / 14.0)]
    tgt = [ctx.array((V,), data=np.sin(x / 3.0) + 0.5 * z / 7.0), ctx.array((V,), data=(x * y) / 49.0)]
    losses = []
    from panslbm2_amd.warmstart import UNet

    nets = [UNet(ctx, dims, 1, 2, seed=5), UNet(ctx, dims, 1, 2, seed=5)]
    for net in nets:
        net.set_input(inp)
        net.set_target(tgt)
        net.forward()
        first = net.backward()
        last = net.train(150, lr=3e-3)
        losses.append((first, last))
    assert losses[0] == losses[1]
    first, last = losses[0]
    print("loss", first, "->", last)
    assert last < 0.1 * first
    out = [ctx.array((V,)) for _ in range(2)]
    nets[0].forward()
    nets[0].output(out)
    for o, t in zip(out, tgt):
        a, b = o.download(), t.download()
        assert np.sqrt(((a - b) ** 2).mean()) < np.sqrt(0.1 * first) + 1e-3
    for net in nets:
        net.close()
