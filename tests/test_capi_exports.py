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
