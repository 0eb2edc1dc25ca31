"""Write the golden fixtures in tests/golden/ from the compiled reference harness.

Run in a container that has the reference tree (to build oracle/_ref) :
    make -C oracle ref && python tests/golden/make_golden.py

For every case in tests/refcases.py it stores the seeded inputs and, per reference function,
the outputs the reference produced.  The fixtures are data only (inputs and expected outputs).
Also stores short driver-sequence runs (lid-driven cavity, coupled forced convection forward and
adjoint) composed from reference calls in the order of test/cavityflow.cpp:44-67 and
test/nsadsens.cpp:67-133, and the output of the C++ heat-sink driver tests/cpp/heatsink_driver.cpp
compiled against the reference headers (oracle/_ref/heatsink_driver_ref).
"""
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import refcases  # noqa: E402
import sequences  # noqa: E402
from _oracle import ref  # noqa: E402


def main(only=None):
    """Write every fixture, or (only = list of case and sequence names) just those fixtures."""
    if ref() is None:
        raise SystemExit("reference harness not built: make -C oracle ref")
    for case in refcases.CASES_2D + refcases.CASES_3D:
        if only and case.name not in only:
            continue
        inp = refcases.make_inputs(case)
        blob = {"in/" + k: np.asarray(v) for k, v in inp.items()}
        for op in refcases.ops_for(case):
            out = refcases.run_ref(case, op, inp)
            for k, v in out.items():
                blob["out/%s/%s" % (op, k)] = np.asarray(v)
        path = os.path.join(HERE, case.name + ".npz")
        np.savez_compressed(path, **blob)
        print("wrote", path, os.path.getsize(path), "bytes")
    for seq in sequences.SEQUENCES:
        if only and seq.name not in only:
            continue
        out = seq.run(sequences.RefBackend())
        path = os.path.join(HERE, "seq_" + seq.name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path), "bytes")
    if not only:
        write_driver_golden()
        write_mma_golden()
        write_heaviside_golden()
        write_vtk_golden()


def write_mma_golden():
    """Fixtures of the reference optimiser (src/utility/mma.h) on the problems of tests/mmacases.py."""
    import ctypes

    import mmacases

    R = ref()
    VP, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    R.refmma_create.restype = VP
    R.refmma_create.argtypes = [I, I, D, VP, VP, VP, VP, VP, D]
    R.refmma_update.argtypes = [VP, I, I, VP, D, VP, VP, VP]
    R.refmma_converged.argtypes = [VP, D]
    R.refmma_destroy.argtypes = [VP]
    p = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(VP)
    for pr in mmacases.PROBLEMS:
        n, m = pr["n"], pr["m"]
        a, c, d = (np.array(pr[k], dtype=np.float64) for k in ("a", "c", "d"))
        lo, hi = mmacases.xbounds(pr)
        h = R.refmma_create(n, m, D(pr["a0"]), p(a), p(c), p(d), p(lo), p(hi), D(pr["move"]))
        x = mmacases.x0(pr)
        blob = {}
        for k in range(pr["steps"]):
            f, dfdx, g, dgdx = mmacases.step_inputs(pr, k, x)
            blob.update({"x_in/%d" % k: x.copy(), "f/%d" % k: np.array(f), "dfdx/%d" % k: dfdx, "g/%d" % k: g,
                         "dgdx/%d" % k: dgdx})
            conv = R.refmma_converged(h, D(f))
            xs = np.ascontiguousarray(x, dtype=np.float64)
            R.refmma_update(h, n, m, xs.ctypes.data_as(VP), D(f), p(dfdx), p(g), p(dgdx))
            blob["x_out/%d" % k] = xs.copy()
            blob["converged/%d" % k] = np.array(conv)
            x = xs
        R.refmma_destroy(h)
        path = os.path.join(HERE, pr["name"] + ".npz")
        np.savez_compressed(path, **blob)
        print("wrote", path, os.path.getsize(path), "bytes")


HEAVISIDE_CASES = [  # (name, lx, ly, lz, R, beta, seed)
    ("h2d_12x10", 12, 10, 1, 1.5, 4.0, 41), ("h3d_7x6x5", 7, 6, 5, 2.4, 8.0, 42), ("h3d_6x5x4_r1", 6, 5, 4, 1.0, 16.0, 43),
]


# (name, lx, ly, lz, R, beta, mx, my, mz, seed): the design-region weights of the drivers
HEAVISIDE_REGION_CASES = [
    ("hr2d_14x11", 14, 11, 1, 2.4, 4.0, 9, 7, 1, 44), ("hr3d_8x7x6", 8, 7, 6, 2.4, 8.0, 5, 6, 4, 45),
]


def heaviside_inputs(lx, ly, lz, seed):
    rng = np.random.default_rng(seed)
    n = lx * ly * lz
    return rng.uniform(0.0, 1.0, n), rng.standard_normal(n)


def write_heaviside_golden():
    """HeavisideFilter (src/utility/heavisidefilter.h, default weights) on the cases above."""
    import ctypes

    R_ = ref()
    VP, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    R_.ref3_heaviside_value.argtypes = [I, I, I, D, D, VP, VP]
    R_.ref3_heaviside_sensitivity.argtypes = [I, I, I, D, D, VP, VP, VP]
    p = lambda a: a.ctypes.data_as(VP)
    blob = {}
    for name, lx, ly, lz, R, beta, seed in HEAVISIDE_CASES:
        s, dfdrho = heaviside_inputs(lx, ly, lz, seed)
        rho, dfds = np.empty_like(s), np.empty_like(s)
        R_.ref3_heaviside_value(lx, ly, lz, D(R), D(beta), p(s), p(rho))
        R_.ref3_heaviside_sensitivity(lx, ly, lz, D(R), D(beta), p(s), p(dfdrho), p(dfds))
        blob.update({name + "/s": s, name + "/dfdrho": dfdrho, name + "/rho": rho, name + "/dfds": dfds})
    R_.ref3_heaviside_value_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP]
    R_.ref3_heaviside_sensitivity_region.argtypes = [I, I, I, D, D, I, I, I, VP, VP, VP]
    for name, lx, ly, lz, R, beta, mx, my, mz, seed in HEAVISIDE_REGION_CASES:
        s, dfdrho = heaviside_inputs(lx, ly, lz, seed)
        rho, dfds = np.empty_like(s), np.empty_like(s)
        R_.ref3_heaviside_value_region(lx, ly, lz, D(R), D(beta), mx, my, mz, p(s), p(rho))
        R_.ref3_heaviside_sensitivity_region(lx, ly, lz, D(R), D(beta), mx, my, mz, p(s), p(dfdrho), p(dfds))
        blob.update({name + "/s": s, name + "/dfdrho": dfdrho, name + "/rho": rho, name + "/dfds": dfds})
    path = os.path.join(HERE, "heaviside.npz")
    np.savez_compressed(path, **blob)
    print("wrote", path, os.path.getsize(path), "bytes")


def write_vtk_golden():
    """The files the reference's VTK writers (and the reader's read-back) produce for
    tests/cpp/vtk_driver.cpp."""
    drv = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "vtk_driver_ref")
    if not os.path.exists(drv):
        raise SystemExit("reference VTK driver not built: make -C oracle ref")
    out = os.path.join(HERE, "vtk")
    os.makedirs(out, exist_ok=True)
    subprocess.run([drv, out], check=True, stdout=subprocess.DEVNULL)
    print("wrote", out, sorted(os.listdir(out)))


def write_driver_golden():
    drv = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "heatsink_driver_ref")
    if not os.path.exists(drv):
        raise SystemExit("reference driver not built: make -C oracle ref")
    path = os.path.join(HERE, "shim_heatsink_30x17.bin")
    tmp = path + ".tmp"
    subprocess.run([drv, tmp], check=True)
    shutil.move(tmp, path)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["driver"]:
        write_driver_golden()
    elif sys.argv[1:] == ["mma"]:
        write_mma_golden()
    elif sys.argv[1:] == ["heaviside"]:
        write_heaviside_golden()
    elif sys.argv[1:] == ["vtk"]:
        write_vtk_golden()
    else:
        main(sys.argv[1:] or None)
