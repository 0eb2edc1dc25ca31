"""ctypes bindings for the test-only CPU oracle (oracle/liblbm_oracle.so) and the reference
harness (oracle/_ref/libref_panslbm2.so, built only where the reference tree exists).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liblbm_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_panslbm2.so")
REF_AVX_SO = os.path.join(ROOT, "oracle", "_ref", "libref_avx.so")

D2Q9, D3Q15, D2Q5, D3Q7, D3Q27 = 0, 1, 2, 3, 4
NC = {D2Q9: 9, D3Q15: 15, D2Q5: 5, D3Q7: 7, D3Q27: 27}
ND = {D2Q9: 2, D3Q15: 3, D2Q5: 2, D3Q7: 3, D3Q27: 3}

VP = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(VP)


def _load(path):
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(ORACLE_SO)
        if _lib is None:
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = _lib
        L.ol_create.restype = VP
        L.ol_create.argtypes = [I, I, I, I]
        L.ol_destroy.argtypes = [VP]
        for n in ("ol_nc", "ol_nxyz", "ol_nface"):
            getattr(L, n).argtypes = [VP]
            getattr(L, n).restype = I
        for n in ("ol_get", "ol_set"):
            getattr(L, n).argtypes = [VP, VP]
        for n in ("ol_get_aos", "ol_set_aos"):
            getattr(L, n).argtypes = [VP, VP, VP]
        L.ol_stream.argtypes = [VP]
        L.ol_istream.argtypes = [VP]
        L.ol_bc.argtypes = [VP, VP, I]
        L.ol_smooth_corner.argtypes = [VP]
        L.ol_ns_init.argtypes = [VP] + [VP] * 4
        L.ol_ns_macro_collide.argtypes = [VP] + [VP] * 4 + [D, I]
        L.ol_ns_macro_brinkman_collide.argtypes = [VP] + [VP] * 4 + [D, VP, I]
        L.ol_ns_set_u.argtypes = [VP] * 5
        L.ol_ns_set_rho.argtypes = [VP] * 5
        L.ol_ad_init.argtypes = [VP] * 5
        L.ol_ad_collide_fc.argtypes = [VP] + [VP] * 5 + [D] + [VP] * 6 + [I, VP, VP]
        L.ol_ad_collide_nc.argtypes = [VP] + [VP] * 5 + [D] + [VP] * 6 + [D] * 4 + [I, VP]
        L.ol_ad_collide_bgk.argtypes = [VP] * 5 + [D, VP] + [VP] * 4 + [D, I] + [D] * 4 + [I]
        L.ol_ad_collide_he.argtypes = [VP] + [VP] * 5 + [D] + [VP] * 6 + [D, I]
        L.ol_ad_set_t.argtypes = [VP] * 6
        L.ol_ad_set_q.argtypes = [VP] * 7 + [D]
        L.ol_ans_init.argtypes = [VP] * 8
        L.ol_ans_collide.argtypes = [VP] + [VP] * 11 + [D, VP, I]
        L.ol_ans_iset_u.argtypes = [VP] * 5 + [D, I]
        L.ol_ans_iset_rho.argtypes = [VP, VP]
        L.ol_nebb_3d.argtypes = [VP, I, VP, VP, VP, VP, D]
        L.ol_aad_init.argtypes = [VP] * 8
        L.ol_aad_collide_fc.argtypes = [VP] + [VP] * 11 + [VP, D, VP] + [VP] * 5 + [VP, I, VP]
        L.ol_aad_collide_nc.argtypes = [VP] + [VP] * 11 + [VP, D, VP] + [VP] * 5 + [VP] + [D] * 3 + [I, VP]
        L.ol_aad_collide_mf.argtypes = [VP] + [VP] * 11 + [VP, D, VP] + [VP] * 5 + [VP] + [D] * 3 + [VP] * 3 + [I, VP]
        L.ol_aad_collide_he.argtypes = [VP] + [VP] * 11 + [VP, D, VP] + [VP] * 5 + [VP, D, I]
        L.ol_aad_iset_t.argtypes = [VP] * 5
        L.ol_aad_iset_q.argtypes = [VP] * 5 + [D]
        L.ol_aad_iset_rho.argtypes = [VP] * 8 + [D]
        L.ol_sens_brinkman.argtypes = [VP] * 9
        L.ol_sens_heat_exchange.argtypes = [VP] * 12
        L.ol_sens_brinkman_diffusivity.argtypes = [VP] * 18
        L.ol_sens_temp_heat_source.argtypes = [VP] * 20
        L.ol_sens_heat_source.argtypes = [VP] * 11
        L.ol_sens_temp_heat_source_c.argtypes = [VP] * 20 + [I]
        L.ol_residual.argtypes = [I] + [VP] * 6 + [I]
        L.ol_residual.restype = D
        L.ol_normalize.argtypes = [VP, I]
        L.ol_density_filter.argtypes = [I, I, I, D, VP, VP]
        L.ol_num_threads.restype = I
    return _lib


_ref = None


def ref():
    """The compiled reference harness, or None where it was not built."""
    global _ref
    if _ref is None:
        _ref = _load(REF_SO)
    return _ref


_ref_avx = None


def ref_avx():
    """The D3Q15 heat-source boundary sensitivity of the reference's AVX build, or None."""
    global _ref_avx
    if _ref_avx is None:
        _ref_avx = _load(REF_AVX_SO)
    return _ref_avx


class Lattice:
    """An oracle lattice: SoA populations f[c, idx] with idx = i + nx*j + nx*ny*k."""

    def __init__(self, kind, lx, ly, lz=1):
        self.kind, self.lx, self.ly = kind, lx, ly
        self.lz = lz if ND[kind] == 3 else 1
        self.nc = NC[kind]
        self.N = lx * ly * self.lz
        self.h = lib().ol_create(kind, lx, ly, self.lz)
        if not self.h:
            raise ValueError("bad lattice")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ol_destroy(self.h)
            self.h = None

    @property
    def nface(self):
        return lib().ol_nface(self.h)

    def get(self):
        a = np.empty((self.nc, self.N))
        lib().ol_get(self.h, ptr(a))
        return a

    def set(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.nc, self.N)
        lib().ol_set(self.h, ptr(a))

    def get_aos(self):
        f0 = np.empty(self.N)
        f = np.empty((self.nc - 1) * self.N)
        lib().ol_get_aos(self.h, ptr(f0), ptr(f))
        return f0, f


def soa_to_aos(soa):
    """SoA (nc, N) -> reference (f0[N], f[(nc-1)*idx + c-1])."""
    soa = np.asarray(soa)
    return soa[0].copy(), np.ascontiguousarray(soa[1:].T).reshape(-1)


def aos_to_soa(f0, f, nc):
    N = f0.shape[0]
    return np.vstack([f0[None, :], np.asarray(f).reshape(N, nc - 1).T])


def face_nodes(kind, lx, ly, lz=1):
    """Node index of every face-array entry (faces xmin, xmax, ymin, ymax[, zmin, zmax])."""
    out = []
    if ND[kind] == 2:
        for face in range(4):
            if face < 2:
                i = 0 if face == 0 else lx - 1
                out += [i + lx * j for j in range(ly)]
            else:
                j = 0 if face == 2 else ly - 1
                out += [i + lx * j for i in range(lx)]
    else:
        for face in range(6):
            ax, hi = divmod(face, 2)
            if ax == 0:
                i = lx - 1 if hi else 0
                out += [i + lx * (t % ly) + lx * ly * (t // ly) for t in range(ly * lz)]
            elif ax == 1:
                j = ly - 1 if hi else 0
                out += [(t % lx) + lx * j + lx * ly * (t // lx) for t in range(lx * lz)]
            else:
                k = lz - 1 if hi else 0
                out += [(t % lx) + lx * (t // lx) + lx * ly * k for t in range(lx * ly)]
    return np.array(out, dtype=np.int64)


def to_face(kind, lx, ly, lz, node_arr, dtype=None):
    """Tabulate a node-indexed array (a driver lambda) on the face-array layout."""
    a = np.asarray(node_arr)[face_nodes(kind, lx, ly, lz)]
    return np.ascontiguousarray(a, dtype=dtype if dtype is not None else a.dtype)
