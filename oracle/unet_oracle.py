"""CPU restatement (numpy, FP64) of the warm-start network -- test infrastructure only.

Only the tests, the smoke check and the CPU leg of the benchmark import this module; the product
path runs the HIP kernels of panslbm2_amd/csrc/unet.hip and never this file.

The network has no reference counterpart.  The reference's TO drivers start every solve from rest
(production/heatsink3D.cpp:97-300 calls NS/AD::InitialCondition with uniform fields, and the
adjoint InitialCondition with zeros); the warm start predicts the fields those InitialCondition
calls take (navierstokes.h:550-572, advection.h:1048-1070, adjointnavierstokes.h:474-498,
adjointadvection.h:1359-1381) from the design.  Parity is pinned here by:
  * conv3d against torch.nn.functional.conv3d in float64 (an independent library routine);
  * the full backward pass against central finite differences of the forward loss;
  * pooling / upsampling against their definitions on brute-force loops.

Layout (shared with the device code):
  * a field on an nx x ny x nz grid is a vector over voxels v = x + nx (y + ny z);
  * activations are [V, C] arrays (channels last);
  * conv weights are [cout, taps, cin] with tap = (dz+1) 9 + (dy+1) 3 + (dx+1) for 3x3x3
    (taps = 27, zero padding 1) and tap = 0 for 1x1x1 (taps = 1); bias [cout];
  * the parameter vector is, conv by conv in NET order, W.ravel() then b.

Architecture (channels c, 2c, 4c with c = base): two 3x3x3 conv + ReLU per level, 2x2x2 average
pooling between levels, nearest 2x upsampling, skip concatenation [upsampled, skip] and a 1x1x1
linear head.  Loss: mean over voxels and output channels of (y - t)^2.

With bf16=True every value the device stores in bfloat16 is rounded the same way here (round to
nearest even): the input, the packed weights, every activation, the pooled and concatenated
tensors, and the gradients handed to the matrix cores; accumulation stays FP64 (FP32 on the
device).
"""
import numpy as np

NET = [  # name, input (tensor name), cin key, cout key, taps
    ("e1a", "x0", "cin", "c1", 27), ("e1b", "e1a", "c1", "c1", 27),
    ("e2a", "p1", "c1", "c2", 27), ("e2b", "e2a", "c2", "c2", 27),
    ("e3a", "p2", "c2", "c3", 27), ("e3b", "e3a", "c3", "c3", 27),
    ("d2a", "cat2", "c32", "c2", 27), ("d2b", "d2a", "c2", "c2", 27),
    ("d1a", "cat1", "c21", "c1", 27), ("d1b", "d1a", "c1", "c1", 27),
    ("head", "d1b", "c1", "cout", 1),
]


def channels(cin, cout, base):
    c1, c2, c3 = base, 2 * base, 4 * base
    return {"cin": cin, "cout": cout, "c1": c1, "c2": c2, "c3": c3, "c32": c3 + c2, "c21": c2 + c1}


def layout(cin, cout, base):
    """[(name, cin, cout, taps, w_offset, b_offset)], total parameter count."""
    ch = channels(cin, cout, base)
    out, off = [], 0
    for name, _, ki, ko, taps in NET:
        a, b = ch[ki], ch[ko]
        out.append((name, a, b, taps, off, off + b * taps * a))
        off += b * taps * a + b
    return out, off


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _grid(x, dims):
    nx, ny, nz = dims
    return x.reshape(nz, ny, nx, -1)


def _taps(taps):
    if taps == 1:
        return [(0, 0, 0)]
    return [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def shifted(x, dims, d):
    """s[v] = x[v + d] with zeros outside the grid."""
    nx, ny, nz = dims
    dx, dy, dz = d
    g = _grid(x, dims)
    p = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
    s = p[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    return s.reshape(nx * ny * nz, -1)


def conv(x, w, b, dims, relu):
    """y[v, o] = b[o] + sum_{tap, i} w[o, tap, i] x[v + off(tap), i]."""
    taps = w.shape[1]
    y = np.tile(b, (x.shape[0], 1)).astype(np.float64)
    for t, d in enumerate(_taps(taps)):
        y += shifted(x, dims, d) @ w[:, t, :].T
    return np.maximum(y, 0.0) if relu else y


def conv_bwd(g, x, w, dims):
    """(dx, dw, db) of conv for the output gradient g."""
    taps = w.shape[1]
    dx = np.zeros_like(x, dtype=np.float64)
    dw = np.zeros_like(w, dtype=np.float64)
    for t, d in enumerate(_taps(taps)):
        dw[:, t, :] = g.T @ shifted(x, dims, d)
        dx += shifted(g @ w[:, t, :], dims, tuple(-e for e in d))
    return dx, dw, g.sum(0)


def pool(x, dims):
    nx, ny, nz = dims
    g = _grid(x, dims)
    c = g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, -1).mean(axis=(1, 3, 5))
    return c.reshape(-1, x.shape[1])


def pool_bwd(gc, dims):
    nx, ny, nz = dims
    c = _grid(gc, (nx // 2, ny // 2, nz // 2)) / 8.0
    f = c.repeat(2, 0).repeat(2, 1).repeat(2, 2)
    return f.reshape(nx * ny * nz, -1)


def up(xc, dims):
    """Nearest upsampling of a coarse field to the fine grid dims."""
    nx, ny, nz = dims
    c = _grid(xc, (nx // 2, ny // 2, nz // 2))
    return c.repeat(2, 0).repeat(2, 1).repeat(2, 2).reshape(nx * ny * nz, -1)


def up_bwd(gf, dims):
    nx, ny, nz = dims
    g = _grid(gf, dims)
    return g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, -1).sum(axis=(1, 3, 5)).reshape(-1, gf.shape[1])


class UNet:
    def __init__(self, dims, cin, cout, base=16):
        nx, ny, nz = dims
        assert nx % 4 == 0 and ny % 4 == 0 and nz % 4 == 0, "two poolings need dims divisible by 4"
        self.dims, self.cin, self.cout, self.base = tuple(dims), cin, cout, base
        self.layers, self.nparams = layout(cin, cout, base)

    def unpack(self, params, bf=False):
        out = {}
        for name, ci, co, taps, wo, bo in self.layers:
            w = params[wo:bo].reshape(co, taps, ci)
            out[name] = (bf16(w) if bf else w, params[bo:bo + co])
        return out

    def forward(self, params, x0, bf=False):
        """Activations dict; 'y' is the head output (FP32 on the device, never rounded)."""
        r = bf16 if bf else (lambda a: a)
        P = self.unpack(params, bf)
        d1 = self.dims
        d2 = tuple(d // 2 for d in d1)
        d3 = tuple(d // 4 for d in d1)
        a = {"x0": r(x0)}
        a["e1a"] = r(conv(a["x0"], *P["e1a"], d1, True))
        a["e1b"] = r(conv(a["e1a"], *P["e1b"], d1, True))
        a["p1"] = r(pool(a["e1b"], d1))
        a["e2a"] = r(conv(a["p1"], *P["e2a"], d2, True))
        a["e2b"] = r(conv(a["e2a"], *P["e2b"], d2, True))
        a["p2"] = r(pool(a["e2b"], d2))
        a["e3a"] = r(conv(a["p2"], *P["e3a"], d3, True))
        a["e3b"] = r(conv(a["e3a"], *P["e3b"], d3, True))
        a["cat2"] = np.concatenate([up(a["e3b"], d2), a["e2b"]], axis=1)
        a["d2a"] = r(conv(a["cat2"], *P["d2a"], d2, True))
        a["d2b"] = r(conv(a["d2a"], *P["d2b"], d2, True))
        a["cat1"] = np.concatenate([up(a["d2b"], d1), a["e1b"]], axis=1)
        a["d1a"] = r(conv(a["cat1"], *P["d1a"], d1, True))
        a["d1b"] = r(conv(a["d1a"], *P["d1b"], d1, True))
        a["y"] = conv(a["d1b"], *P["head"], d1, False)
        return a

    def loss(self, params, x0, target, bf=False):
        y = self.forward(params, x0, bf)["y"]
        return float(((y - target) ** 2).mean())

    def grad(self, params, x0, target, bf=False):
        """(loss, gradient vector in the parameter layout)."""
        r = bf16 if bf else (lambda a: a)
        P = self.unpack(params, bf)
        a = self.forward(params, x0, bf)
        d1 = self.dims
        d2 = tuple(d // 2 for d in d1)
        d3 = tuple(d // 4 for d in d1)
        y = a["y"]
        loss = float(((y - target) ** 2).mean())
        G = {}
        relu = lambda g, act: r(g * (act > 0))

        g = r(2.0 * (y - target) / y.size)
        dx, *G["head"] = conv_bwd(g, a["d1b"], P["head"][0], d1)
        g = relu(dx, a["d1b"])
        dx, *G["d1b"] = conv_bwd(g, a["d1a"], P["d1b"][0], d1)
        g = relu(dx, a["d1a"])
        dx, *G["d1a"] = conv_bwd(g, a["cat1"], P["d1a"][0], d1)
        c2 = 2 * self.base
        skip1 = dx[:, c2:]
        g = relu(up_bwd(dx[:, :c2], d1), a["d2b"])
        dx, *G["d2b"] = conv_bwd(g, a["d2a"], P["d2b"][0], d2)
        g = relu(dx, a["d2a"])
        dx, *G["d2a"] = conv_bwd(g, a["cat2"], P["d2a"][0], d2)
        c3 = 4 * self.base
        skip2 = dx[:, c3:]
        g = relu(up_bwd(dx[:, :c3], d2), a["e3b"])
        dx, *G["e3b"] = conv_bwd(g, a["e3a"], P["e3b"][0], d3)
        g = relu(dx, a["e3a"])
        dx, *G["e3a"] = conv_bwd(g, a["p2"], P["e3a"][0], d3)
        g = relu(skip2 + pool_bwd(dx, d2), a["e2b"])
        dx, *G["e2b"] = conv_bwd(g, a["e2a"], P["e2b"][0], d2)
        g = relu(dx, a["e2a"])
        dx, *G["e2a"] = conv_bwd(g, a["p1"], P["e2a"][0], d2)
        g = relu(skip1 + pool_bwd(dx, d1), a["e1b"])
        dx, *G["e1b"] = conv_bwd(g, a["e1a"], P["e1b"][0], d1)
        g = relu(dx, a["e1a"])
        _, *G["e1a"] = conv_bwd(g, a["x0"], P["e1a"][0], d1)

        out = np.zeros(self.nparams)
        for name, ci, co, taps, wo, bo in self.layers:
            dw, db = G[name]
            out[wo:bo] = dw.ravel()
            out[bo:bo + co] = db
        return loss, out


def adam(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1), in place; t counts from 1."""
    m[:] = b1 * m + (1 - b1) * grad
    v[:] = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    params -= lr * mh / (np.sqrt(vh) + eps)
