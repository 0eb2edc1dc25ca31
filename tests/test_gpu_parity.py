"""HIP path against the reference's outputs (golden fixtures written from the compiled reference).

Bar: stream, lattice boundaries, corner smoothing, every collide, initial conditions, boundary
operators, Normalize, the density filter and whole driver sequences are value-identical to the
reference (FP64, no FMA contraction; +0 == -0).  Functions whose reference arithmetic calls glibc
pow() or sums in a different order -- Residual (residual.h) and the pow-based sensitivity terms
(adjointadvection.h:1473, :766) -- are held to a relative 1e-12.
"""
import os

import numpy as np
import pytest

import refcases
import sequences

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REL_TOL = {"residual": 1e-12, "sens_brinkman_diffusivity": 1e-12, "sens_temp_heat_source": 1e-12}


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _compare(got, want, what, rtol=None):
    got, want = np.asarray(got), np.asarray(want)
    assert got.shape == want.shape, what
    assert np.all(np.isfinite(got)) or not np.all(np.isfinite(want)), what + ": non-finite values"
    if rtol is None:
        bad = got != want
        assert not bad.any(), "%s: %d values differ (max abs %.3g)" % (what, int(bad.sum()),
                                                                       float(np.max(np.abs(got - want))))
    else:
        scale = max(float(np.max(np.abs(want))), 1e-300)
        err = float(np.max(np.abs(got - want))) / scale
        assert err <= rtol, "%s: relative error %.3g > %.1g" % (what, err, rtol)


def _cases():
    return [(c, op) for c in refcases.CASES_2D + refcases.CASES_3D for op in refcases.ops_for(c)]


@pytest.mark.parametrize("case,op", _cases(), ids=lambda x: getattr(x, "name", x))
def test_gpu_function_matches_reference(ctx, case, op):
    from gpu_cases import run_gpu

    print("case", case.name, op, flush=True)
    blob = np.load(os.path.join(GOLDEN, case.name + ".npz"))
    inp = {k[3:]: blob[k] for k in blob.files if k.startswith("in/")}
    pre = "out/%s/" % op
    want = {k[len(pre):]: blob[k] for k in blob.files if k.startswith(pre)}
    got = run_gpu(ctx, case, op, inp)
    for k, v in want.items():
        _compare(got[k], v, "%s %s %s" % (case.name, op, k), REL_TOL.get(op))


@pytest.mark.parametrize("fused", [False, True], ids=["eager", "fused"])
@pytest.mark.parametrize("seq", sequences.SEQUENCES, ids=lambda s: s.name)
def test_gpu_sequence_matches_reference(ctx, seq, fused):
    from gpu_backend import FusedGpuBackend, GpuBackend

    print("sequence", seq.name, "fused" if fused else "eager", flush=True)
    b = FusedGpuBackend(ctx) if fused else GpuBackend(ctx)
    got = seq.run(b)
    if fused:
        assert b.fused_calls > 0
    blob = np.load(os.path.join(GOLDEN, "seq_" + seq.name + ".npz"))
    for k in blob.files:
        tol = 1e-12 if k in ("residuals", "dfds") else None
        _compare(got[k], blob[k], "%s %s" % (seq.name, k), tol)
