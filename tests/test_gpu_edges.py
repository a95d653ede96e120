"""Edge cases through the C ABI on the GPU: an empty mesh, non-finite query
points, and the error codes of undersized buffers and calls out of order
(include/its_c.h: ITS_E_INVALID -1, ITS_E_STATE -4, ITS_E_RANGE -5).

Non-finite points: the reference hat (BSpline.hpp:5-27) returns 0 outside
its support, so +-inf and huge coordinates give the outside value in the
support query (mq_setInDomainPoints, SVO.cu:1195-1261) and 0 in the dense
sum.  A NaN coordinate passes the reference's box test (no comparison is
true) and then fails every comparison of its hat, which runs off the end of
the function: undefined there.  The library's hat weights are
max(0, 1 - |t|), which is 0 for a NaN t, so such a point lies outside every
support and f = 0 in every mode (the oracle's brute-force hat gives NaN;
neither is the reference's).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import fitted

pytestmark = pytest.mark.gpu

E_INVALID, E_STATE, E_RANGE = -1, -4, -5


def _gpu(ctx, F):
    from its_amd import _capi
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)
    g.set_lambda(F.lam)
    return g


def test_empty_mesh(ctx):
    from its_amd import _capi
    origin, width = np.zeros(3), np.ones(3)
    print("[edges] empty mesh, R=16", flush=True)
    g = _capi.Octree(ctx, np.zeros((0, 9)), origin, width, 16)
    ref = O.SVO(np.zeros((0, 9)), origin, width, 16)
    assert (g.depth, g.num_nodes, g.num_verts, g.num_voxels) == (ref.depth, ref.num_nodes, ref.num_verts,
                                                                  ref.num_voxels) == (0, 0, 0, 0)
    # every later step reports the missing octree instead of failing on the device
    # (a zero-length host array may also be passed as a null pointer)
    with pytest.raises(_capi.ItsError) as e:
        g.lscg(np.zeros(0), tol=1e-6)
    assert e.value.code in (E_STATE, E_INVALID)
    with pytest.raises(_capi.ItsError) as e:
        g.set_lambda(np.zeros(0))
    assert e.value.code in (E_STATE, E_INVALID)
    with pytest.raises(_capi.ItsError) as e:
        g.eval(np.zeros((4, 3), np.float32), origin + width)
    assert e.value.code == E_STATE
    with pytest.raises(_capi.ItsError) as e:
        g.marching_cubes(origin, width, (8, 8, 8), 0.0)
    assert e.value.code == E_STATE
    g.close()


@pytest.mark.parametrize("mode", ["support", "dense"])
def test_non_finite_queries(ctx, mode):
    from its_amd import _capi
    F = fitted("icosphere", 16)
    c = F.case
    print(f"[edges] non-finite queries, {mode}", flush=True)
    g = _gpu(ctx, F)
    big = [[np.inf, 0.5, 0.5], [0.5, -np.inf, 0.5], [0.5, 0.5, 1e30], [-1e30, 0.5, 0.5], [0.5, 1e30, -np.inf]]
    nan = [[np.nan, 0.5, 0.5], [0.5, 0.5, np.nan], [np.nan, np.nan, np.nan]]
    finite = [[0.3, 0.4, 0.5], [0.61, 0.52, 0.47]]
    P = np.array(big + nan + finite)
    out_val = 0.25
    if mode == "support":
        f = g.eval(P.astype(np.float32), c.end, _capi.EVAL_SUPPORT, np.float32(out_val)).astype(np.float64)
        ref = F.svo.eval_support(F.lam, P.astype(np.float32).astype(np.float64), c.end, out_val)
    else:
        f = g.eval(P.astype(np.float32), c.end, _capi.EVAL_DENSE).astype(np.float64)
        ref = F.svo.eval_dense(F.lam, P.astype(np.float32).astype(np.float64))
    nb, nn = len(big), len(nan)
    expect_big = out_val if mode == "support" else 0.0
    assert np.all(f[:nb] == expect_big) and np.all(ref[:nb] == expect_big)
    assert np.all(f[nb:nb + nn] == 0.0)
    assert np.abs(f[nb + nn:] - ref[nb + nn:]).max() <= 1e-5


def test_error_codes(ctx):
    from its_amd import _capi
    F = fitted("icosphere", 16)
    c = F.case
    print("[edges] error codes", flush=True)
    g = _gpu(ctx, F)
    L = _capi.lib()
    res = np.array([16, 16, 16], np.int32)
    st = _capi.McStats()
    iso = np.zeros(1)
    nv = np.zeros(1, np.int64)
    assert L.its_mc_extract(ctx.handle, g.handle, None, c.origin, c.width, res, iso, 1, None, 0, nv,
                            C.byref(st)) == 0
    n = st.num_verts
    assert n > 3 and nv[0] == n
    buf = np.zeros((n, 3), np.float32)
    # one vertex short: ITS_E_RANGE, nothing written
    rc = L.its_mc_extract(ctx.handle, g.handle, None, c.origin, c.width, res, iso, 1, buf.ctypes.data_as(C.c_void_p),
                          n - 1, nv, C.byref(st))
    assert rc == E_RANGE and not buf.any()
    # a resolution of zero: ITS_E_INVALID
    rc = L.its_mc_extract(ctx.handle, g.handle, None, c.origin, c.width, np.array([16, 0, 16], np.int32), iso, 1,
                          None, 0, nv, C.byref(st))
    assert rc == E_INVALID
    # no iso value, or more than ITS_MC_MAX_ISO: ITS_E_INVALID
    for k in (0, 17):
        rc = L.its_mc_extract(ctx.handle, g.handle, None, c.origin, c.width, res, np.zeros(max(k, 1)), k, None, 0,
                              np.zeros(max(k, 1), np.int64), C.byref(st))
        assert rc == E_INVALID, k
    # the same call afterwards still works
    tri, st2 = g.marching_cubes(c.origin, c.width, (16, 16, 16), 0.0)
    assert st2["num_verts"] == n
    # R out of range: ITS_E_RANGE
    with pytest.raises(_capi.ItsError) as e:
        _capi.Octree(ctx, c.tris, c.origin, c.width, 12)
    assert e.value.code == E_RANGE
    # evaluation before any coefficients: ITS_E_STATE
    g2 = _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)
    with pytest.raises(_capi.ItsError) as e:
        g2.eval(np.zeros((4, 3), np.float32), c.end)
    assert e.value.code == E_STATE
    g2.close()
    g.close()
