"""The 3D heat-sink driver (panslbm2_amd/heatsink3d.py): its fused steps (one kernel pass per step)
against the same boundary calls run one by one (each value-identical to the reference, see the
3D sequence tests) -- FP64, value-identical -- and its mixed-precision evaluation against FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = dict(lx=14, ly=10, lz=9, plate=3)


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _state(h):
    d = h.d
    names = ("rho", "tem", "ux0", "ux1", "uz0", "qz1", "ip", "iux0", "iuz1", "imz", "item", "iqz0")
    return [h.pf.download(), h.pg.download()] + [d[k].download() for k in names]


@pytest.mark.parametrize("ns_kind", ["D3Q15", "D3Q27"])
def test_fused_steps_equal_eager_steps(ctx, ns_kind):
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    k = getattr(pa, ns_kind)
    hs = {f: Heatsink3D(ctx, **DIMS, seed=4, use_graphs=False, fused=f, max_steps=100, ns_kind=k)
          for f in (True, False)}
    for h in hs.values():
        h.init_forward()
        for t in range(1, 121):
            h._fwd_step(t)
        h.fwd_buf = 0
        h.init_adjoint()
        for t in range(1, 61):
            h._adj_step(t, 0)
    for x, y in zip(_state(hs[True]), _state(hs[False])):
        np.testing.assert_array_equal(x, y)
    for h in hs.values():
        h.close()


@pytest.mark.parametrize("ns_kind", ["D3Q15", "D3Q27"])
def test_mixed_precision_evaluation(ctx, ns_kind):
    """FP32 storage (switching to FP64 near the end) against FP64, both solved to 1e-9 so that only
    precision separates them: objective 1e-5, dfds 1e-4 of its largest entry."""
    import panslbm2_amd as pa
    from panslbm2_amd.heatsink3d import Heatsink3D

    out = {}
    for prec in (pa.F64, pa.F32):
        h = Heatsink3D(ctx, **DIMS, seed=4, prec=prec, eps=1e-9, max_steps=200000, ns_kind=getattr(pa, ns_kind))
        r = h.evaluate()
        print("3d prec", prec, r["fwd_steps"], r["adj_steps"], r["fp32_steps"], r["objective"], flush=True)
        assert not r["diverged"] and max(r["fwd_residual"]) < 1e-9 and max(r["adj_residual"]) < 1e-9
        out[prec] = (r["objective"], h.d["dfds"].download())
        h.close()
    (o64, s64), (o32, s32) = out[pa.F64], out[pa.F32]
    assert abs(o32 - o64) <= 1e-5 * abs(o64)
    assert np.max(np.abs(s32 - s64)) <= 1e-4 * np.max(np.abs(s64))
