"""The MMA optimiser on the device (lbm_mma_* in include/panslbm2_amd.h), after the reference's
MMA class (src/utility/mma.h:59-142, UpdateVariables :196-548)."""
import ctypes

import numpy as np

from . import VP, check


def _host(a, n):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))


class MMA:
    """n design variables on the device, m constraints g_i(x) <= 0.  update() takes device arrays
    (DeviceArray) for x (replaced by the new design), dfdx and dgdx (m x n), and host values f, g."""

    def __init__(self, ctx, n, m, a0, a, c, d, xmin, xmax, move=None):
        self.ctx, self.L = ctx, ctx.lib
        self.n, self.m = n, m
        a, c, d = (_host(v, m) for v in (a, c, d))
        lo = xmin if hasattr(xmin, "ptr") else ctx.array((n,), data=_host(xmin, n))
        hi = xmax if hasattr(xmax, "ptr") else ctx.array((n,), data=_host(xmax, n))
        h = VP()
        check(self.L.lbm_mma_create(ctx.h, n, m, a0, a.ctypes.data_as(VP), c.ctypes.data_as(VP),
                                    d.ctypes.data_as(VP), lo.ptr, hi.ptr, ctypes.byref(h)))
        self.h = h
        self.iterations = 0
        if move is not None:
            self.set("move", move)

    def set(self, name, value):
        check(self.L.lbm_mma_set(self.h, name.encode(), float(value)))

    def converged(self, f):
        out = ctypes.c_int()
        check(self.L.lbm_mma_converged(self.h, float(f), ctypes.byref(out)))
        return bool(out.value)

    def update(self, x, f, dfdx, g, dgdx):
        g = _host(g, self.m)
        it = ctypes.c_int()
        check(self.L.lbm_mma_update(self.h, x.ptr, float(f), dfdx.ptr, g.ctypes.data_as(VP), dgdx.ptr,
                                    ctypes.byref(it)))
        self.iterations = it.value
        return it.value

    def close(self):
        if getattr(self, "h", None):
            self.L.lbm_mma_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
