"""Learned warm start of the steady solves (lbm_unet_* in include/panslbm2_amd.h).

UNet wraps one network on the device: a 3-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions,
bfloat16 weights and activations, FP32 accumulation on the matrix cores) that maps input fields
(the design and boundary masks) to output fields (the macroscopic fields the InitialCondition
calls take).  WarmStart is the online predictor a design loop uses: it keeps the last few
converged solutions, trains the network on them with Adam after every solve, and writes its
prediction for the next design into the solver's arrays.  The solver's residual still decides
convergence; a warm start that diverges is replaced by the cold start (Heatsink3D.evaluate).

No reference counterpart: the reference starts every solve from uniform fields
(production/heatsink3D.cpp:97-300).
"""
import ctypes

import numpy as np

from . import SZ, VP, check


def _ptrs(arrays):
    return (ctypes.c_void_p * len(arrays))(*[a.ptr.value if isinstance(a.ptr, VP) else a.ptr for a in arrays])


def _dbl(v, n):
    if v is None:
        return None
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
    return a


class UNet:
    def __init__(self, ctx, dims, cin, cout, base=16, seed=0):
        self.ctx, self.L = ctx, ctx.lib
        self.dims = tuple(dims)
        self.cin, self.cout = cin, cout
        h = VP()
        nx, ny, nz = self.dims
        check(self.L.lbm_unet_create(ctx.h, nx, ny, nz, cin, cout, base, seed, ctypes.byref(h)))
        self.h = h
        n = SZ()
        check(self.L.lbm_unet_num_params(h, ctypes.byref(n)))
        self.nparams = n.value
        self.V = nx * ny * nz

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.nparams)
        check(self.L.lbm_unet_get_params(self.h, out.ctypes.data_as(VP), self.nparams))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        check(self.L.lbm_unet_set_params(self.h, p.ctypes.data_as(VP), p.size))

    def grads(self):
        out = np.empty(self.nparams)
        check(self.L.lbm_unet_get_grads(self.h, out.ctypes.data_as(VP), self.nparams))
        return out

    def reset_optimizer(self):
        check(self.L.lbm_unet_reset_optimizer(self.h))

    def _fields(self, fn, arrays, mean, scale):
        n = len(arrays)
        m, s = _dbl(mean, n), _dbl(scale, n)
        check(fn(self.h, _ptrs(arrays), None if m is None else m.ctypes.data_as(VP),
                 None if s is None else s.ctypes.data_as(VP), n))

    def set_input(self, arrays, mean=None, scale=None):
        self._fields(self.L.lbm_unet_set_input, arrays, mean, scale)

    def set_target(self, arrays, mean=None, scale=None):
        self._fields(self.L.lbm_unet_set_target, arrays, mean, scale)

    def forward(self):
        check(self.L.lbm_unet_forward(self.h))

    def output(self, arrays, mean=None, scale=None):
        self._fields(self.L.lbm_unet_output, arrays, mean, scale)

    def backward(self):
        loss = ctypes.c_double()
        check(self.L.lbm_unet_backward(self.h, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(self.L.lbm_unet_adam(self.h, lr, beta1, beta2, eps))

    def train(self, steps, lr=1e-3):
        loss = ctypes.c_double()
        check(self.L.lbm_unet_train(self.h, steps, lr, ctypes.byref(loss)))
        return loss.value

    def read(self, name, channels):
        """[voxels (of the tensor's level), channels] host copy of a named tensor."""
        lvl = {"p1": 1, "e2a": 1, "e2b": 1, "cat2": 1, "d2a": 1, "d2b": 1, "p2": 2, "e3a": 2, "e3b": 2}.get(name, 0)
        n = (self.V >> (3 * lvl)) * channels
        out = np.empty(n)
        check(self.L.lbm_unet_read(self.h, name.encode(), out.ctypes.data_as(VP), n))
        return out.reshape(-1, channels)


class WarmStart:
    """Online predictor of converged fields.

    observe(inputs, targets) stores a copy of one converged solution (device arrays) and trains
    `steps` Adam updates over the stored solutions (the newest `replay` of them, round robin).
    predict(inputs, outputs) writes the prediction for the given inputs into `outputs`.
    The per-field normalisation (mean, scale) is fixed from the first observed solution.
    """

    def __init__(self, ctx, dims, n_in, n_out, seed=0, lr=1e-3, steps=40, replay=4, min_samples=1):
        self.ctx = ctx
        self.net = UNet(ctx, dims, n_in, n_out, seed=seed)
        self.lr, self.steps, self.replay, self.min_samples = lr, steps, replay, min_samples
        self.samples = []  # [(inputs copies, targets copies)]
        self.mean = self.scale = None
        self.last_loss = None
        self.n_observed = 0

    @property
    def ready(self):
        return self.n_observed >= self.min_samples

    def observe(self, inputs, targets):
        if self.mean is None:
            host = [t.download().ravel() for t in targets]
            self.mean = np.array([h.mean() for h in host])
            self.scale = np.array([max(h.std(), 1e-12) for h in host])
        if len(self.samples) >= self.replay:
            ins, outs = self.samples.pop(0)
        else:
            ins = [self.ctx.array(a.shape) for a in inputs]
            outs = [self.ctx.array(a.shape) for a in targets]
        for dst, src in zip(ins + outs, list(inputs) + list(targets)):
            check(self.ctx.lib.lbm_device_copy(self.ctx.h, dst.ptr, src.ptr, 8 * src.size))
        self.samples.append((ins, outs))
        self.n_observed += 1
        loss = None
        for s in range(self.steps):
            ins, outs = self.samples[(len(self.samples) - 1 - s) % len(self.samples)]
            self.net.set_input(ins)
            self.net.set_target(outs, self.mean, self.scale)
            self.net.forward()
            if s == self.steps - 1:
                loss = self.net.backward()  # waits for the device once per observation
            else:
                check(self.net.L.lbm_unet_backward(self.net.h, None))
            self.net.adam(self.lr)
        self.last_loss = loss
        return loss

    def predict(self, inputs, outputs):
        self.net.set_input(inputs)
        self.net.forward()
        self.net.output(outputs, self.mean, self.scale)

    def close(self):
        self.net.close()
