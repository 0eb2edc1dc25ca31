"""The C-ABI library loads (no GPU needed) and exports every function include/panslbm2_amd.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "panslbm2_amd.h")
LIB = os.path.join(ROOT, "panslbm2_amd", "libpanslbm2_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) > 50
    for must in ("lbm_stream", "lbm_ad_macro_brinkman_collide_force_convection", "lbm_residual",
                 "lbm_sensitivity_temperature_at_heat_source", "lbm_record_begin"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, "not exported: %s" % missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_reports_errors_without_gpu():
    import panslbm2_amd as pa

    L = pa.load()
    assert L.lbm_version() >= 1
    h = ctypes.c_void_p()
    rc = L.lbm_context_create(10 ** 6, ctypes.byref(h))
    assert rc != 0 and L.lbm_last_error()
    # entry points that take a context refuse a null one with an argument error
    assert L.lbm_context_set_concurrent_smoothing(None, 1) == -1  # LBM_ERR_ARG
    assert b"null" in L.lbm_last_error()


def declared_params():
    """name -> list of parameter declarations, for every function the header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(lbm_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        out[m.group(1)] = params
    return out


def _ctype_of(param):
    import panslbm2_amd as pa

    if "*" in param:
        return pa.VP
    base = param.rsplit(None, 1)[0].replace("const ", "").strip()
    return {"double": pa.D, "int": pa.I, "size_t": pa.SZ, "float": ctypes.c_float,
            "unsigned long long": ctypes.c_ulonglong}.get(base, pa.VP)


def _kind(t):
    if t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer):
        return "pointer"
    return t.__name__


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_bindings_match_header():
    """Every ctypes signature the Python package sets has the header's parameter count and kinds
    (pointer / double / int / size_t), so a call through it cannot shift its arguments."""
    import panslbm2_amd as pa

    L = pa.load()
    decl = declared_params()
    checked = 0
    for name, params in decl.items():
        fn = getattr(L, name)
        if fn.argtypes is None:
            continue
        want = [_kind(_ctype_of(p)) for p in params]
        got = [_kind(t) for t in fn.argtypes]
        assert got == want, "%s: bound %s, header %s" % (name, fn.argtypes, params)
        checked += 1
    assert checked > 50
