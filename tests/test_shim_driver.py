"""The templated C++ drop-in (include/panslbm2_amd.hpp) driven by a reference-style C++ driver.

tests/cpp/heatsink_driver.cpp is written against the reference's API (D2Q9<double>, NS::, AD::,
ANS::, AAD::, Residual, Normalize) in the call order of test/nsadsens.cpp:67-133 and
production/heatsink.cpp:140-244: two design iterations of forward solve, adjoint solve and three
sensitivity reductions, with the drivers' velocity-array swaps.  The golden file holds what the
same source printed when compiled against the reference headers; the GPU build must leave the
same values in every host array.

Tolerances: populations and macroscopic fields are compared exactly (the device path keeps the
reference's operation order without FMA contraction).  Sensitivities use pow(x, 2.0) in the
reference where the device squares by multiplication, and Residual sums in a different order, so
those compare to 1e-12 relative to the array's largest magnitude.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden", "shim_heatsink_30x17.bin")
DRIVER = os.path.join(HERE, "cpp", "bin", "heatsink_driver")
LOOSE = ("dfds", "dfds2", "dfdsn", "residuals")


def read_records(path):
    out = {}
    b = open(path, "rb").read()
    o = 0
    while o < len(b):
        name = b[o:o + 32].split(b"\0")[0].decode()
        o += 32
        n = int(np.frombuffer(b[o:o + 8], np.int64)[0])
        o += 8
        out[name] = np.frombuffer(b[o:o + 8 * n], np.float64).copy()
        o += 8 * n
    return out


def test_golden_is_complete():
    d = read_records(GOLDEN)
    fields = ["f", "rho", "ux", "tem", "g", "iux", "item", "ig", "dfds", "dfds2", "dfdsn", "pf_f", "pg_f"]
    for design in (0, 1):
        for k in fields:
            v = d["d%d_%s" % (design, k)]
            assert v.size and np.isfinite(v).all(), k
        # the run moved away from the initial state and the normalised sensitivity peaks at 1
        assert np.abs(d["d%d_ux" % design]).max() > 1e-3
        assert np.abs(d["d%d_dfdsn" % design]).max() == pytest.approx(1.0)
    assert d["residuals"].size == 2 * 4 * (240 // 40)


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "panslbm2_amd.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


THERMAL_USE = r"""
#include "panslbm2_amd.hpp"
using namespace PANSLBM2;
void use2(D2Q9<double> &p, D2Q5<double> &q, double *a) {
    AD::MacroBrinkmanCollideNaturalConvection(p, a, a, a, a, 0.1, q, a, a, a, a, 0.0, -1e-4, 0.5, true, a);
    AD::MacroBrinkmanCollideHeatExchange(p, a, a, a, a, 0.1, q, a, a, a, a, 0.05, true);
    AAD::MacroBrinkmanCollideNaturalConvection(p, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, 0.0, -1e-4, true, a);
    AAD::MacroBrinkmanCollideHeatExchange(p, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, 0.05, true);
    AAD::SensitivityHeatExchange(q, a, a, a, a, a, a, a, a, a);
    AD::MacroCollideForceConvection(p, a, a, a, 0.02, q, a, a, a, 0.02, true);
    AD::MacroCollideNaturalConvection(p, a, a, a, 0.02, q, a, a, a, 0.02, 0.0, 1.6e-5, 1.5, true);
    AAD::MacroBrinkmanCollideNaturalConvectionMassFlow(p, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, 0.0, -1e-4,
                                                       a, a, true, a);
}
void use3(D3Q15<double> &p, D3Q7<double> &q, double *a) {
    AD::MacroBrinkmanCollideNaturalConvection(p, a, a, a, a, a, 0.1, q, a, a, a, a, a, 0.0, 0.0, -1e-4, 0.5, true, a);
    AD::MacroBrinkmanCollideHeatExchange(p, a, a, a, a, a, 0.1, q, a, a, a, a, a, 0.05, true);
    AAD::MacroBrinkmanCollideNaturalConvection(p, a, a, a, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, a, 0.0,
                                               0.0, -1e-4, true, a);
    AAD::MacroBrinkmanCollideHeatExchange(p, a, a, a, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, a, 0.05, true);
    AAD::SensitivityHeatExchange(q, a, a, a, a, a, a, a, a, a, a, a);
    AD::MacroCollideForceConvection(p, a, a, a, a, 0.02, q, a, a, a, a, 0.02, true);
    AD::MacroCollideNaturalConvection(p, a, a, a, a, 0.02, q, a, a, a, a, 0.02, 0.0, 0.0, 1.6e-5, 1.5, true);
    AAD::MacroBrinkmanCollideNaturalConvectionMassFlow(p, a, a, a, a, a, a, a, a, a, a, a, a, 0.1, q, a, a, a, a, a, a,
                                                       0.0, 0.0, -1e-4, a, a, a, true, a);
}
void use27(D3Q27<double> &p, D3Q7<double> &q, double *a) {
    NS::MacroBrinkmanCollide(p, a, a, a, a, 0.1, a, true);
    AD::MacroBrinkmanCollideForceConvection(p, a, a, a, a, a, 0.1, q, a, a, a, a, a, true);
    p.Stream();
    p.iStream();
    p.SmoothCorner();
}
int main() { return 0; }
"""


def test_thermal_variants_instantiate(tmp_path):
    """The natural-convection and heat-exchange overloads take the reference's argument lists
    (advection.h:528-712, :716-810 and :930-1044, adjointadvection.h:894-993, :1113-1358 and
    :1445-1457; the callers test/naturalconvection.cpp:48 and production/ncpump.cpp:193):
    calls written in those orders compile against the header; so do the 3D calls on D3Q27."""
    src = tmp_path / "thermal.cpp"
    src.write_text(THERMAL_USE)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_driver_built():
    assert os.path.exists(DRIVER), "build the C++ drivers first (python -c 'import __graft_entry__ as g; g.build()')"


def compare(got, want):
    assert set(got) == set(want)
    for k, w in want.items():
        g = got[k]
        assert g.shape == w.shape, k
        if k.split("_", 1)[-1] in LOOSE:
            scale = max(np.abs(w).max(), 1e-300)
            err = np.abs(g - w).max() / scale
            assert err <= 1e-12, "%s: relative error %.3e" % (k, err)
        else:
            bad = np.flatnonzero(g != w)
            assert bad.size == 0, "%s: %d entries differ, first at %d: %r vs %r" % (
                k, bad.size, bad[0], g[bad[0]], w[bad[0]])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["host", "device"])
def test_driver_matches_reference(tmp_path, mode):
    out = tmp_path / ("driver_%s.bin" % mode)
    print("heatsink_driver", mode, flush=True)
    r = subprocess.run([DRIVER, str(out), mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    compare(read_records(str(out)), read_records(GOLDEN))
