"""lbm_record_end_deferred: a recording ended without building its device tables.  It compares
equal to the same recording ended by lbm_record_end, applying it gives the same populations (the
tables are built at its first use), and a first use inside a graph capture is refused with
LBM_ERR_STATE while the stream stays usable."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VP = ctypes.c_void_p


@pytest.fixture(scope="module")
def ctx():
    import panslbm2_amd as pa

    c = pa.Context(0)
    yield c
    c.close()


def _record(ctx, lat, walls, ux, uy, zero, inlet, deferred):
    import panslbm2_amd as pa

    L = ctx.lib
    pa.check(L.lbm_record_begin(ctx.h))
    pa.check(L.lbm_stream(lat.h))
    pa.check(L.lbm_bc(lat.h, pa.hptr(walls)))
    pa.check(L.lbm_ns_set_u(lat.h, pa.hptr(inlet), pa.hptr(ux), pa.hptr(zero), pa.hptr(zero)))
    pa.check(L.lbm_smooth_corner(lat.h))
    prog = VP()
    end = L.lbm_record_end_deferred if deferred else L.lbm_record_end
    pa.check(end(ctx.h, ctypes.byref(prog)))
    return prog


def test_deferred_program_equals_and_applies_like_the_uploaded_one(ctx):
    import panslbm2_amd as pa
    from panslbm2_amd import D2Q9, to_face

    lx, ly = 14, 10
    N = lx * ly
    n = np.arange(N)
    i, j = n % lx, n // lx
    walls = to_face(D2Q9, lx, ly, 1, ((j == 0) | (j == ly - 1)).astype(np.int32), np.int32)
    inlet = to_face(D2Q9, lx, ly, 1, ((i == 0) & (j > 0) & (j < ly - 1)).astype(np.int32), np.int32)
    ux = to_face(D2Q9, lx, ly, 1, np.full(N, 0.03))
    zero = to_face(D2Q9, lx, ly, 1, np.zeros(N))
    rng = np.random.default_rng(3)
    f0 = rng.uniform(0.01, 0.2, (9, N))
    out = []
    progs = []
    for deferred in (False, True):
        lat = pa.Lattice(ctx, D2Q9, lx, ly)
        lat.upload(f0)
        prog = _record(ctx, lat, walls, ux, zero, zero, inlet, deferred)
        progs.append((lat, prog))
        pa.check(ctx.lib.lbm_program_apply(prog))
        out.append(lat.download())
    np.testing.assert_array_equal(out[0], out[1])
    # a third recording on the first lattice, deferred, equals the first one and is dropped unused
    lat0, p0 = progs[0]
    p2 = _record(ctx, lat0, walls, ux, zero, zero, inlet, True)
    assert ctx.lib.lbm_program_equal(p0, p2) == 1
    pa.check(ctx.lib.lbm_program_destroy(p2))
    for lat, prog in progs:
        pa.check(ctx.lib.lbm_program_destroy(prog))
        lat.close()


def test_deferred_program_first_used_in_a_capture_is_refused(ctx):
    import panslbm2_amd as pa
    from panslbm2_amd import D2Q9, to_face

    lx, ly = 12, 9
    N = lx * ly
    n = np.arange(N)
    j = n // lx
    walls = to_face(D2Q9, lx, ly, 1, ((j == 0) | (j == ly - 1)).astype(np.int32), np.int32)
    zero = to_face(D2Q9, lx, ly, 1, np.zeros(N))
    none = to_face(D2Q9, lx, ly, 1, np.zeros(N, dtype=np.int32), np.int32)
    lat = pa.Lattice(ctx, D2Q9, lx, ly)
    lat.upload(np.full((9, N), 0.1))
    prog = _record(ctx, lat, walls, zero, zero, zero, none, True)
    L = ctx.lib
    pa.check(L.lbm_graph_begin(ctx.h))
    rc = L.lbm_program_apply(prog)
    msg = L.lbm_last_error().decode()
    g = VP()
    rc_end = L.lbm_graph_end(ctx.h, ctypes.byref(g))
    if rc_end == 0 and g:
        L.lbm_graph_destroy(g)
    assert rc == -3  # LBM_ERR_STATE
    assert "graph capture" in msg
    # outside the capture the same program uploads and runs
    pa.check(L.lbm_program_apply(prog))
    assert np.all(np.isfinite(lat.download()))
    pa.check(L.lbm_program_destroy(prog))
    lat.close()
