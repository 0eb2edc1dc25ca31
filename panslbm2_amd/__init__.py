"""panslbm2_amd -- MI355X (gfx950) HIP implementation of the PANSLBM2 lattice Boltzmann hot path.

The product is the C-ABI library panslbm2_amd/libpanslbm2_hip.so (include/panslbm2_amd.h).  This
module is a thin ctypes binding used by Python callers, the tests and bench.py.  It never falls
back to anything else: if the library is missing, load() raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpanslbm2_hip.so")

D2Q9, D3Q15, D2Q5, D3Q7, D3Q27 = 0, 1, 2, 3, 4
NC = {D2Q9: 9, D3Q15: 15, D2Q5: 5, D3Q7: 7, D3Q27: 27}
ND = {D2Q9: 2, D3Q15: 3, D2Q5: 2, D3Q7: 3, D3Q27: 3}
F64, F32, F16 = 0, 1, 2  # population storage (lbm_lattice_create_prec)
BARRIER, MIRROR = 1, 2
SET_T, SET_Q = 1, 2

VP = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
SZ = ctypes.c_size_t

_lib = None


class LBMError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LBMError("%s is missing: build it with `python -m panslbm2_amd.build`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.lbm_last_error.restype = ctypes.c_char_p
    sig = {
        "lbm_heaviside_filter": [VP, I, I, I, D, D, VP, VP],
        "lbm_heaviside_sensitivity": [VP, I, I, I, D, D, VP, VP, VP],
        "lbm_heaviside_filter_region": [VP, I, I, I, D, D, I, I, I, VP, VP],
        "lbm_heaviside_sensitivity_region": [VP, I, I, I, D, D, I, I, I, VP, VP, VP],
        "lbm_mma_create": [VP, I, I, D, VP, VP, VP, VP, VP, ctypes.POINTER(VP)],
        "lbm_mma_destroy": [VP],
        "lbm_mma_set": [VP, ctypes.c_char_p, D],
        "lbm_mma_update": [VP, VP, D, VP, VP, VP, ctypes.POINTER(I)],
        "lbm_mma_converged": [VP, D, ctypes.POINTER(I)],
        "lbm_unet_create": [VP, I, I, I, I, I, I, ctypes.c_ulonglong, ctypes.POINTER(VP)],
        "lbm_unet_destroy": [VP],
        "lbm_unet_num_params": [VP, ctypes.POINTER(SZ)],
        "lbm_unet_get_params": [VP, VP, SZ],
        "lbm_unet_set_params": [VP, VP, SZ],
        "lbm_unet_get_grads": [VP, VP, SZ],
        "lbm_unet_reset_optimizer": [VP],
        "lbm_unet_set_input": [VP, VP, VP, VP, I],
        "lbm_unet_set_target": [VP, VP, VP, VP, VP, I],
        "lbm_unet_synchronize": [VP],
        "lbm_unet_forward": [VP],
        "lbm_unet_output": [VP, VP, VP, VP, VP, VP, VP, I],
        "lbm_unet_backward": [VP, VP],
        "lbm_unet_adam": [VP, D, D, D, D],
        "lbm_unet_train": [VP, I, D, VP],
        "lbm_unet_read": [VP, ctypes.c_char_p, VP, SZ],
        "lbm_context_create": [I, ctypes.POINTER(VP)],
        "lbm_context_destroy": [VP],
        "lbm_context_synchronize": [VP],
        "lbm_context_set_concurrent_smoothing": [VP, I],
        "lbm_device_alloc": [VP, SZ, ctypes.POINTER(VP)],
        "lbm_device_free": [VP, VP],
        "lbm_copy_to_device": [VP, VP, VP, SZ],
        "lbm_copy_to_host": [VP, VP, VP, SZ],
        "lbm_device_fill": [VP, VP, D, SZ],
        "lbm_device_copy": [VP, VP, VP, SZ],
        "lbm_event_record": [VP, I],
        "lbm_event_elapsed_ms": [VP, I, I, ctypes.POINTER(ctypes.c_float)],
        "lbm_lattice_create": [VP, I, I, I, I, ctypes.POINTER(VP)],
        "lbm_lattice_create_prec": [VP, I, I, I, I, I, ctypes.POINTER(VP)],
        "lbm_lattice_precision": [VP],
        "lbm_lattice_set_precision": [VP, I],
        "lbm_batch_begin": [VP],
        "lbm_batch_end": [VP, ctypes.POINTER(VP)],
        "lbm_batch_size": [VP],
        "lbm_batch_launch": [VP],
        "lbm_batch_set_active": [VP, VP],
        "lbm_batch_configure": [VP, I, I],
        "lbm_batch_tune": [ctypes.POINTER(VP), I, I, ctypes.POINTER(I), ctypes.POINTER(I)],
        "lbm_batch_configure_tiles": [VP, I, I, I],
        "lbm_batch_tune_tiles": [ctypes.POINTER(VP), I, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I),
                                 ctypes.POINTER(I)],
        "lbm_batch_destroy": [VP],
        "lbm_residual_batch": [VP, I, I, ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(I),
                               ctypes.POINTER(D)],
        "lbm_residual_batch_static": [VP, I, I, ctypes.POINTER(VP), ctypes.POINTER(VP), ctypes.POINTER(I),
                                      ctypes.POINTER(D)],
        "lbm_lattice_destroy": [VP],
        "lbm_lattice_nc": [VP],
        "lbm_lattice_nxyz": [VP],
        "lbm_lattice_nface": [VP],
        "lbm_lattice_upload": [VP, VP],
        "lbm_lattice_download": [VP, VP],
        "lbm_lattice_upload_aos": [VP, VP, VP],
        "lbm_lattice_download_aos": [VP, VP, VP],
        "lbm_stream": [VP],
        "lbm_istream": [VP],
        "lbm_bc": [VP, VP],
        "lbm_ibc": [VP, VP],
        "lbm_smooth_corner": [VP],
        "lbm_ns_initial_condition": [VP] * 5,
        "lbm_ns_macro_collide": [VP] * 5 + [D, I],
        "lbm_ns_macro_brinkman_collide": [VP] * 5 + [D, VP, I],
        "lbm_ns_set_u": [VP] * 5,
        "lbm_ns_set_rho": [VP] * 5,
        "lbm_ad_initial_condition": [VP] * 5,
        "lbm_ad_macro_brinkman_collide_force_convection": [VP] * 6 + [D] + [VP] * 6 + [I, VP, VP],
        "lbm_ad_macro_brinkman_collide_natural_convection": [VP] * 6 + [D] + [VP] * 6 + [D] * 4 + [I, VP],
        "lbm_ad_macro_brinkman_collide_heat_exchange": [VP] * 6 + [D] + [VP] * 6 + [D, I],
        "lbm_ad_macro_collide_force_convection": [VP] * 5 + [D] + [VP] * 5 + [D, I],
        "lbm_ad_macro_collide_force_convection_p": [VP] * 6 + [D] + [VP] * 5 + [D, I],
        "lbm_ad_macro_collide_natural_convection": [VP] * 5 + [D] + [VP] * 5 + [D] * 5 + [I],
        "lbm_ad_macro_collide_natural_convection_p": [VP] * 6 + [D] + [VP] * 5 + [D] * 5 + [I],
        "lbm_aad_macro_brinkman_collide_natural_convection_mass_flow":
            [VP] * 13 + [D] + [VP] * 7 + [D] * 3 + [VP] * 3 + [I, VP],
        "lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p":
            [VP] * 14 + [D] + [VP] * 7 + [D] * 3 + [VP] * 3 + [I, VP],
        "lbm_ad_macro_brinkman_collide_heat_exchange_p": [VP] * 7 + [D] + [VP] * 6 + [D, I],
        "lbm_aad_macro_brinkman_collide_heat_exchange": [VP] * 13 + [D] + [VP] * 7 + [D, I],
        "lbm_aad_macro_brinkman_collide_heat_exchange_p": [VP] * 14 + [D] + [VP] * 7 + [D, I],
        "lbm_sensitivity_heat_exchange": [VP] * 12,
        "lbm_sensitivity_heat_source": [VP] * 11,
        "lbm_ad_set_t": [VP] * 6,
        "lbm_ad_set_q": [VP] * 7 + [D],
        "lbm_ans_initial_condition": [VP] * 8,
        "lbm_ans_macro_brinkman_collide": [VP] * 12 + [D, VP, I],
        "lbm_ans_iset_u": [VP] * 5 + [D, I],
        "lbm_ans_iset_rho": [VP, VP],
        "lbm_aad_initial_condition": [VP] * 8,
        "lbm_aad_macro_brinkman_collide_force_convection": [VP] * 13 + [D] + [VP] * 7 + [I, VP],
        "lbm_aad_macro_brinkman_collide_natural_convection": [VP] * 13 + [D] + [VP] * 7 + [D] * 3 + [I, VP],
        "lbm_aad_iset_t": [VP] * 5,
        "lbm_aad_iset_q": [VP] * 5 + [D],
        "lbm_aad_iset_rho": [VP] * 8 + [D],
        "lbm_sensitivity_brinkman": [VP] * 9,
        "lbm_sensitivity_brinkman_diffusivity": [VP] * 18,
        "lbm_sensitivity_temperature_at_heat_source": [VP] * 20,
        "lbm_sensitivity_temperature_at_heat_source_c": [VP] * 20 + [I],
        "lbm_residual": [VP, I] + [VP] * 6 + [I, ctypes.POINTER(D)],
        "lbm_normalize": [VP, VP, I],
        "lbm_gather_sum": [VP, VP, VP, I, ctypes.POINTER(D)],
        "lbm_density_filter": [VP, I, I, I, D, VP, VP],
        "lbm_record_begin": [VP],
        "lbm_record_end": [VP, ctypes.POINTER(VP)],
        "lbm_record_end_deferred": [VP, ctypes.POINTER(VP)],
        "lbm_program_destroy": [VP],
        "lbm_program_apply": [VP],
        "lbm_program_bind_fields": [VP, I] + [VP] * 6,
        "lbm_program_num_field_sets": [VP],
        "lbm_program_fusable": [VP],
        "lbm_program_equal": [VP, VP],
        "lbm_graph_begin": [VP],
        "lbm_graph_end": [VP, ctypes.POINTER(VP)],
        "lbm_graph_launch": [VP],
        "lbm_graph_destroy": [VP],
        "lbm_ns_macro_collide_p": [VP] * 6 + [D, I],
        "lbm_ns_macro_brinkman_collide_p": [VP] * 6 + [D, VP, I],
        "lbm_ad_macro_brinkman_collide_force_convection_p": [VP] * 7 + [D] + [VP] * 6 + [I, VP, VP],
        "lbm_ad_macro_brinkman_collide_natural_convection_p": [VP] * 7 + [D] + [VP] * 6 + [D] * 4 + [I, VP],
        "lbm_ans_macro_brinkman_collide_p": [VP] * 13 + [D, VP, I],
        "lbm_aad_macro_brinkman_collide_force_convection_p": [VP] * 14 + [D] + [VP] * 7 + [I, VP],
        "lbm_aad_macro_brinkman_collide_natural_convection_p": [VP] * 14 + [D] + [VP] * 7 + [D] * 3 + [I, VP],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.lbm_lattice_populations.argtypes = [VP]
    L.lbm_lattice_populations.restype = VP
    L.lbm_version.restype = I
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise LBMError("libpanslbm2_hip error %d: %s" % (rc, load().lbm_last_error().decode()))
    return rc


def hptr(a):
    """ctypes pointer to a C-contiguous host numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(VP)


class Context:
    """One device and one HIP stream (lbm_context)."""

    def __init__(self, device=0):
        self.lib = load()
        h = VP()
        check(self.lib.lbm_context_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.lib.lbm_context_synchronize(self.h))

    def array(self, shape, fill=None, data=None):
        return DeviceArray(self, shape, fill=fill, data=data)

    def record(self, event):
        check(self.lib.lbm_event_record(self.h, event))

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        check(self.lib.lbm_event_elapsed_ms(self.h, a, b, ctypes.byref(ms)))
        return ms.value


class DeviceArray:
    """A float64 device allocation (lbm_device_alloc)."""

    def __init__(self, ctx, shape, fill=None, data=None):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.size = int(np.prod(self.shape))
        p = VP()
        check(ctx.lib.lbm_device_alloc(ctx.h, 8 * self.size, ctypes.byref(p)))
        self.ptr = p
        if data is not None:
            self.upload(data)
        elif fill is not None:
            self.fill(fill)

    @property
    def p(self):
        return self.ptr

    def fill(self, v):
        check(self.ctx.lib.lbm_device_fill(self.ctx.h, self.ptr, float(v), self.size))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.size, (a.shape, self.shape)
        check(self.ctx.lib.lbm_copy_to_device(self.ctx.h, self.ptr, hptr(a), 8 * self.size))

    def download(self):
        a = np.empty(self.shape, dtype=np.float64)
        check(self.ctx.lib.lbm_copy_to_host(self.ctx.h, hptr(a), self.ptr, 8 * self.size))
        return a

    def free(self):
        if self.ptr:
            self.ctx.lib.lbm_device_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Lattice:
    """Device populations of one lattice (lbm_lattice)."""

    def __init__(self, ctx, kind, lx, ly, lz=1, prec=None):
        self.ctx, self.kind, self.lx, self.ly = ctx, kind, lx, ly
        self.lz = lz if ND[kind] == 3 else 1
        self.nc = NC[kind]
        self.N = lx * ly * self.lz
        self.prec = F64 if prec is None else prec
        h = VP()
        check(ctx.lib.lbm_lattice_create_prec(ctx.h, kind, lx, ly, self.lz, self.prec, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.lbm_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.close()
        except Exception:
            pass

    @property
    def nface(self):
        return self.ctx.lib.lbm_lattice_nface(self.h)

    def upload(self, soa):
        a = np.ascontiguousarray(soa, dtype=np.float64).reshape(self.nc, self.N)
        check(self.ctx.lib.lbm_lattice_upload(self.h, hptr(a)))

    def download(self):
        a = np.empty((self.nc, self.N))
        check(self.ctx.lib.lbm_lattice_download(self.h, hptr(a)))
        return a


def face_nodes(kind, lx, ly, lz=1):
    """Node index of each face-array entry (faces xmin, xmax, ymin, ymax[, zmin, zmax])."""
    if ND[kind] == 2:
        j = np.arange(ly)
        i = np.arange(lx)
        return np.concatenate([0 + lx * j, (lx - 1) + lx * j, i, i + lx * (ly - 1)]).astype(np.int64)
    t_x = np.arange(ly * lz)
    t_y = np.arange(lx * lz)
    t_z = np.arange(lx * ly)
    xf = lambda i: i + lx * (t_x % ly) + lx * ly * (t_x // ly)
    yf = lambda j: (t_y % lx) + lx * j + lx * ly * (t_y // lx)
    zf = lambda k: (t_z % lx) + lx * (t_z // lx) + lx * ly * k
    return np.concatenate([xf(0), xf(lx - 1), yf(0), yf(ly - 1), zf(0), zf(lz - 1)]).astype(np.int64)


def to_face(kind, lx, ly, lz, node_arr, dtype=np.float64):
    """Tabulate a node-indexed array (a reference lambda) on the face-array layout."""
    return np.ascontiguousarray(np.asarray(node_arr)[face_nodes(kind, lx, ly, lz)], dtype=dtype)
