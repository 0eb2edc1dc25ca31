"""lbm_residual_batch_static, the no-allocation form of lbm_residual_batch: identical results on two
calls in a row (the second with fewer and shorter array sets, so the shared scratch holds the first
call's contents when the second starts), the size limit, and the same results from a second
context on the same GPU (one scratch per GPU)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _sets(ctx, rng, lens, ncomp):
    n = len(lens)
    arrs = [[ctx.array((m,), data=rng.standard_normal(m)) for _ in range(2 * ncomp)] for m in lens]
    a = (ctypes.c_void_p * (3 * n))()
    b = (ctypes.c_void_p * (3 * n))()
    for i, x in enumerate(arrs):
        for k in range(ncomp):
            a[3 * i + k] = x[k].ptr
            b[3 * i + k] = x[ncomp + k].ptr
    return arrs, a, b, (ctypes.c_int * n)(*lens)


def _run(ctx, fn, n, ncomp, a, b, lens):
    import panslbm2_amd as pa

    out = np.full(n, np.nan)
    pa.check(fn(ctx.h, n, ncomp, a, b, lens, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def test_static_residual_equals_allocating_form(ctx):
    rng = np.random.default_rng(17)
    L = ctx.lib
    cases = [([3000, 257, 100000, 33, 4096, 1, 70000, 640] * 4, 3),  # 32 sets, 3 components
             ([513, 20], 2)]                                          # then fewer, shorter sets
    for lens, ncomp in cases:
        print("static residual: %d sets, ncomp %d" % (len(lens), ncomp), flush=True)
        arrs, a, b, cl = _sets(ctx, rng, lens, ncomp)
        want = _run(ctx, L.lbm_residual_batch, len(lens), ncomp, a, b, cl)
        got = _run(ctx, L.lbm_residual_batch_static, len(lens), ncomp, a, b, cl)
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, want)


def test_static_residual_limit_and_second_context(ctx):
    import panslbm2_amd as pa

    L = ctx.lib
    rng = np.random.default_rng(18)
    n = pa.LBM_RESIDUAL_STATIC_MAX if hasattr(pa, "LBM_RESIDUAL_STATIC_MAX") else 256
    lens = [64] * (n + 1)
    arrs, a, b, cl = _sets(ctx, rng, lens, 1)
    out = np.zeros(n + 1)
    rc = L.lbm_residual_batch_static(ctx.h, n + 1, 1, a, b, cl, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == -1  # LBM_ERR_ARG: beyond the scratch's capacity, nothing launched
    assert b"at most" in L.lbm_last_error()
    # exactly at the limit, and from a second context on the same device
    want = _run(ctx, L.lbm_residual_batch, n, 1, a, b, cl)
    c2 = pa.Context(0)
    try:
        got = _run(c2, L.lbm_residual_batch_static, n, 1, a, b, cl)
    finally:
        c2.close()
    np.testing.assert_array_equal(got, want)
