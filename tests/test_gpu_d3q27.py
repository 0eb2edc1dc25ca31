"""D3Q27 Navier-Stokes (no reference counterpart) coupled with D3Q7 advection-diffusion on the HIP
path, against the oracle run on the same seeded inputs in this process.

The oracle's D3Q27 operators are pinned in tests/test_oracle_d3q27.py (equality of the
construction with the reference's D3Q15 operators, moment constraints, adjoint-transpose checks).
Bar, as for the reference lattices: FP64 values identical (+0 == -0) for every function, except
the functions whose oracle arithmetic calls pow() (Residual and the pow-based sensitivity terms),
held to 1e-12 relative; FP32 storage to 1e-5 of each array's largest magnitude.
"""
import numpy as np
import pytest

import refcases

pytestmark = pytest.mark.gpu

REL_TOL = {"residual": 1e-12, "sens_brinkman_diffusivity": 1e-12, "sens_temp_heat_source": 1e-12}
FP32_SKIP = ("residual", "normalize", "density_filter", "sens_brinkman", "sens_brinkman_diffusivity",
             "sens_temp_heat_source", "sens_heat_exchange")
COPY_OPS = ("stream", "istream", "bc", "ibc")


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _cases(skip=()):
    return [(c, op) for c in refcases.CASES_D3Q27 for op in refcases.ops_for(c) if op not in skip]


@pytest.mark.parametrize("case,op", _cases(), ids=lambda x: getattr(x, "name", x))
def test_d3q27_function_matches_oracle(ctx, case, op):
    from gpu_cases import run_gpu

    print("d3q27 case", case.name, op, flush=True)
    inp = refcases.make_inputs(case)
    want = refcases.run_oracle(case, op, inp)
    got = run_gpu(ctx, case, op, inp)
    for k, v in want.items():
        g, w = np.asarray(got[k]), np.asarray(v)
        assert g.shape == w.shape, k
        rtol = REL_TOL.get(op)
        if rtol is None:
            bad = g != w
            assert not bad.any(), "%s %s %s: %d values differ (max abs %.3g)" % (
                case.name, op, k, int(bad.sum()), float(np.max(np.abs(g - w))))
        else:
            err = float(np.max(np.abs(g - w))) / max(float(np.max(np.abs(w))), 1e-300)
            assert err <= rtol, "%s %s %s: relative error %.3g" % (case.name, op, k, err)


@pytest.mark.parametrize("case,op", _cases(FP32_SKIP), ids=lambda x: getattr(x, "name", x))
def test_d3q27_fp32_matches_oracle(ctx, case, op):
    import panslbm2_amd as pa
    from gpu_cases import run_gpu

    print("d3q27 fp32 case", case.name, op, flush=True)
    inp = refcases.make_inputs(case)
    want = refcases.run_oracle(case, op, inp)
    got = run_gpu(ctx, case, op, inp, prec=pa.F32)
    for k, v in want.items():
        g, w = np.asarray(got[k], dtype=np.float64), np.asarray(v, dtype=np.float64)
        if op in COPY_OPS and k in ("f", "g"):
            np.testing.assert_array_equal(g, w.astype(np.float32).astype(np.float64))
            continue
        assert np.all(np.isfinite(g))
        err = float(np.max(np.abs(g - w))) / max(float(np.max(np.abs(w))), 1e-300)
        assert err <= 1e-5, "%s %s %s: relative error %.3g" % (case.name, op, k, err)


@pytest.mark.parametrize("concurrent", [1, 0], ids=["concurrent", "after_pass"])
@pytest.mark.parametrize("kinds", [("D3Q27", "D3Q7"), ("D3Q15", "D3Q7")], ids=["q27-q7", "q15-q7"])
def test_fused_smoothed_sequence_matches_eager(ctx, kinds, concurrent):
    """The 3D forced-convection sequence with SmoothCorner on both lattices, forward and adjoint
    (tests/sequences.py ForcedConvection3D) on NS + D3Q7 lattice pairs: the fused steps (the pass,
    and the smoothed edges and corners either concurrently with it, from face states each lane
    rebuilds, or after it, from the face states the pass saved) value-identical to the eager
    sequence (stream, boundary passes, smoothing, collide).  D3Q27 + D3Q7 fills 34 of the side
    state's components."""
    import _oracle
    import panslbm2_amd as pa
    import sequences
    from gpu_backend import FusedGpuBackend, GpuBackend

    k = tuple(getattr(_oracle, n) for n in kinds)
    print("fused smoothed sequence", kinds, "concurrent" if concurrent else "after the pass", flush=True)
    seq = sequences.ForcedConvection3D("fc3d_%s_%s" % kinds, 7, 5, 5, 30, check=15, kinds=k)
    want = seq.run(GpuBackend(ctx))
    fb = FusedGpuBackend(ctx)
    pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, concurrent))
    try:
        got = seq.run(fb)
    finally:
        pa.check(ctx.lib.lbm_context_set_concurrent_smoothing(ctx.h, 1))
    assert fb.fused_calls > 0
    for key, w in want.items():
        np.testing.assert_array_equal(np.asarray(got[key]), np.asarray(w), err_msg=key)
