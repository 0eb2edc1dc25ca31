"""Learned warm start of the steady solves (lbm_unet_* in include/panslbm2_amd.h).

UNet wraps one network on the device: a 3-level 3D U-Net (16/32/64 channels, 3x3x3 convolutions,
bfloat16 weights and activations, FP32 accumulation on the matrix cores) that maps input fields
(the design, its change, and coordinates) to output fields (the macroscopic fields the
InitialCondition calls take).  WarmStart is the online predictor a design loop uses: the network
learns the change of the converged fields from one design to the next, trained with Adam on the
solutions the loop produces; its prediction for a new design is the previous converged fields plus
that learned change.  Training runs on the network's own stream, beside the solver.  The solver's
residual still decides convergence; a warm start that diverges is replaced by the cold start
(Heatsink3D.evaluate).

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

    def synchronize(self):
        check(self.L.lbm_unet_synchronize(self.h))

    @staticmethod
    def _vec(v, n):
        a = _dbl(v, n)
        return None if a is None else a.ctypes.data_as(VP), a

    def set_input(self, arrays, mean=None, scale=None):
        n = len(arrays)
        (m, _m), (s, _s) = self._vec(mean, n), self._vec(scale, n)
        check(self.L.lbm_unet_set_input(self.h, _ptrs(arrays), m, s, n))

    def set_target(self, arrays, base=None, mean=None, scale=None):
        n = len(arrays)
        (m, _m), (s, _s) = self._vec(mean, n), self._vec(scale, n)
        check(self.L.lbm_unet_set_target(self.h, _ptrs(arrays), None if base is None else _ptrs(base), m, s, n))

    def forward(self):
        check(self.L.lbm_unet_forward(self.h))

    def output(self, arrays, base=None, mean=None, scale=None, lo=None, hi=None):
        n = len(arrays)
        v = [self._vec(a, n) for a in (mean, scale, lo, hi)]
        check(self.L.lbm_unet_output(self.h, _ptrs(arrays), None if base is None else _ptrs(base),
                                     *[p for p, _ in v], n))

    def backward(self, want_loss=True):
        """Loss of the last forward pass (waits for the network stream), or None with
        want_loss=False (nothing waits)."""
        if not want_loss:
            check(self.L.lbm_unet_backward(self.h, None))
            return None
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
    """Online predictor of converged fields, as the change from the previous converged solution.

    observe(inputs, targets): after a converged solve, stores (inputs, targets, previous targets)
        as a training sample and keeps a copy of targets as the base of the next prediction.
    train(): queues `steps` Adam updates over the newest `replay` samples (round robin) on the
        network's stream and returns at once; the solve that follows runs beside it.
    predict(inputs, outputs): outputs = previous targets + gain x predicted change, each field
        clamped to the range its observed values spanned (widened by 5 % on either side).  Before
        any training the network outputs zero, so the first warm start is the previous solution.
    gain: per field, the least-squares scale of the last prediction's change onto the change the
        solve then found, <p, t> / <p, p> clipped to [0, 1] -- a held-out check, since the network
        had not trained on that design.  A network whose change does not point the right way is
        thereby shrunk toward the previous solution; one that does keeps gain 1.
    The normalisation of the change (per-field mean and scale) is fixed from the first sample.
    learn=False never trains: every prediction is then the previous solution (the plain
    continuation warm start the learned one is measured against).
    """

    GAIN_MAX = 1.0

    def __init__(self, ctx, dims, n_in, n_out, seed=0, lr=1e-3, steps=40, replay=4, learn=True):
        self.ctx = ctx
        self.net = UNet(ctx, dims, n_in, n_out, seed=seed)
        self.lr, self.steps, self.replay, self.learn = lr, steps, replay, learn
        self.samples = []  # [(inputs, targets, base)] device copies
        self.prev = None   # copies of the last observed targets
        self.mean = np.zeros(n_out)
        self.scale = np.ones(n_out)
        self.have_stats = False
        self.lo = self.hi = None
        self.n_observed = 0
        self.trained_steps = 0
        self.gain = np.ones(n_out)
        self.prev_host = None  # host copies of the last observed targets
        self.raw = None        # host copy of the last prediction's change at gain 1

    @property
    def ready(self):
        return self.prev is not None

    def _copies(self, src, dst=None):
        if dst is None:
            dst = [self.ctx.array(a.shape) for a in src]
        for d, a in zip(dst, src):
            check(self.ctx.lib.lbm_device_copy(self.ctx.h, d.ptr, a.ptr, 8 * a.size))
        return dst

    def observe(self, inputs, targets):
        host = [t.download().ravel() for t in targets]
        if self.raw is not None and self.prev_host is not None:
            for f, (p, h, b) in enumerate(zip(self.raw, host, self.prev_host)):
                pp = float(np.dot(p, p))
                self.gain[f] = min(max(float(np.dot(p, h - b)) / pp, 0.0), self.GAIN_MAX) if pp > 0.0 else 1.0
        self.raw = None
        self.prev_host = host
        lo = np.array([h.min() for h in host])
        hi = np.array([h.max() for h in host])
        self.lo = lo if self.lo is None else np.minimum(self.lo, lo)
        self.hi = hi if self.hi is None else np.maximum(self.hi, hi)
        if self.prev is not None and self.learn:
            if not self.have_stats:
                base = [b.download().ravel() for b in self.prev]
                self.mean = np.array([(h - b).mean() for h, b in zip(host, base)])
                self.scale = np.array([max((h - b).std(), 1e-30) for h, b in zip(host, base)])
                self.have_stats = True
            old = self.samples.pop(0) if len(self.samples) >= self.replay else (None, None, None)
            ins = self._copies(inputs, old[0])
            outs = self._copies(targets, old[1])
            base = self._copies(self.prev, old[2])
            self.samples.append((ins, outs, base))
        self.prev = self._copies(targets, self.prev)
        self.n_observed += 1

    def train(self, steps=None):
        steps = self.steps if steps is None else steps
        if not self.samples or not self.learn:
            return
        for s in range(steps):
            ins, outs, base = self.samples[(len(self.samples) - 1 - s) % len(self.samples)]
            self.net.set_input(ins)
            self.net.set_target(outs, base, self.mean, self.scale)
            self.net.forward()
            self.net.backward(want_loss=False)
            self.net.adam(self.lr)
        self.trained_steps += steps

    def sample_loss(self, k=-1):
        """Loss on stored sample k (waits for the device); None without samples."""
        if not self.samples:
            return None
        ins, outs, base = self.samples[k]
        self.net.set_input(ins)
        self.net.set_target(outs, base, self.mean, self.scale)
        self.net.forward()
        return self.net.backward()

    def predict(self, inputs, outputs):
        self.net.set_input(inputs)
        self.net.forward()
        pad = 0.05 * (self.hi - self.lo)
        lo, hi = self.lo - pad, self.hi + pad
        if not self.learn or not self.samples:  # untrained: the previous solution
            self.net.output(outputs, self.prev, self.mean, self.scale, lo, hi)
            return
        self.net.output(outputs, self.prev, self.mean, self.scale, lo, hi)
        self.raw = [o.download().ravel() - b for o, b in zip(outputs, self.prev_host)]
        if np.any(self.gain != 1.0):
            self.net.output(outputs, self.prev, self.mean * self.gain, self.scale * self.gain, lo, hi)

    def close(self):
        self.net.close()
