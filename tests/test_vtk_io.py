"""The VTK writers and reader of the C++ drop-in (include/panslbm2_amd_vtk.hpp) against the
reference's (src/utility/vtkexport.h, vtkxmlexport.h, vtkimport.h).

tests/cpp/vtk_driver.cpp makes the calls the reference's drivers make: legacy files in 2D and 3D
with scalars (double, int and float values), vectors and tensors; XML .pvts + .vts pairs with 1,
3 and 9 components, several arrays per file, and one with no point data; then reads the 3D legacy
file back through VTKImport (every GetPointScalar / GetPointVector / GetPointTensor overload).
Built against the drop-in header, its files must equal byte for byte the ones the same source
wrote through the reference's headers (tests/golden/vtk, make_golden.py vtk), and so must its
standard output (the reader reports each array it finds).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden", "vtk")
FILES = ["legacy2d.vtk", "legacy3d.vtk", "xml2d.pvts", "xml2d_0.vts", "xml3d.pvts", "xml3d_0.vts", "xmlempty.pvts",
         "xmlempty_0.vts", "import.txt"]


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("vtk")
    exe = tmp / "vtk_driver"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", str(exe), os.path.join(HERE, "cpp", "vtk_driver.cpp")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    out = tmp / "out"
    out.mkdir()
    r = subprocess.run([str(exe), str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    return out, r.stdout


@pytest.mark.parametrize("name", FILES)
def test_file_matches_reference_bytes(written, name):
    out, _ = written
    got = (out / name).read_bytes()
    with open(os.path.join(GOLDEN, name), "rb") as f:
        want = f.read()
    assert got == want, "%s differs from the reference's file (first difference at byte %d)" % (
        name, next((i for i, (a, b) in enumerate(zip(got, want)) if a != b), min(len(got), len(want))))


def test_reader_reports_what_it_finds(written):
    _, stdout = written
    assert stdout.splitlines() == ['Find out "ss"', 'Find out "q"', 'Find out "sigma"', 'Find out "sigma"',
                                   'Find out "q"']


def test_reader_round_trip_is_the_written_values(written):
    """import.txt holds what the reader parsed; the writer printed 6 significant digits, so the
    parsed scalars agree with the legacy file's text exactly and the columns are laid out as read."""
    out, _ = written
    lines = (out / "import.txt").read_text().splitlines()
    nx, ny, nz = map(int, lines[0].split())
    assert (nx, ny, nz) == (5, 3, 4) and len(lines) == 1 + nx * ny * nz
    text = (out / "legacy3d.vtk").read_text().split("\n")
    start = text.index("LOOKUP_TABLE\tdefault") + 1
    ss = [float(v) for v in text[start:start + nx * ny * nz]]
    got = [float(line.split()[0]) for line in lines[1:]]
    assert got == ss
