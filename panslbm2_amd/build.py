"""Build the HIP library (gfx950) and the test-only CPU oracle.

    python -m panslbm2_amd.build            # library + oracle (+ reference harness if present)
                                            # + the C++ drivers of tests/cpp

The library is built in-tree (panslbm2_amd/libpanslbm2_hip.so) so that it travels with the tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpanslbm2_hip.so")
BUILD = os.path.join(HERE, "_build")
# heaviest first: the 3D coupled collides, then the rest, so the parallel build ends together
SOURCES = ["collide_%s%s%s.hip" % (eq, nd, p) for nd in ("3", "2") for eq in ("ad", "aad", "ns", "ans")
           for p in ("", "_f32", "_f16")] + ["device.hip", "capi.hip", "unet.hip", "mma.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no FMA contraction, so FP64 kernels reproduce the reference's plain
# x86-64 arithmetic.  Packed FP32 instructions are disabled for all device code.
# -simplifycfg-sink-common=false: the per-face boundary operators write different populations;
# sinking those stores into one store through a selected address forces the node's populations
# into scratch memory.  Without the sinking they stay in registers.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
         "-mllvm", "-simplifycfg-sink-common=false"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(min(len(SOURCES), max(2, os.cpu_count() or 2))) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=False):
    out = _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src"):
        out += _run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    if verbose:
        print(out)


def build_drivers(verbose=False):
    """C++ drivers written against include/panslbm2_amd.hpp (tests/cpp), linked to the library."""
    out = _run(["make", "-B", "-C", os.path.join(ROOT, "tests", "cpp")])
    if verbose:
        print(out)


def main():
    build_oracle(verbose=True)
    build_library(verbose=True)
    build_drivers(verbose=True)


if __name__ == "__main__":
    main()
