"""B-spline evaluation on the GPU (binary32) against the FP64 oracle.

Reference paths: the CPU query mt_cpuTest_2 (ThinShells.cpp:1249-1294), the
GPU dense row sums (CUDACompute.cu:435-590, 1070-1156) and
cpLatentBSplineValue (ThinShells.cpp:655-701).  Tolerance: 1e-5 absolute
on f, which is the bound stated for f(x) against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
from helpers import case, fitted

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _gpu(ctx, F):
    from its_amd import _capi
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)
    g.set_lambda(F.lam)
    return g


def _queries(c, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = c.origin - 0.05, c.end + 0.05
    return rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)


@pytest.mark.parametrize("name,R", [("icosphere", 64), ("bunny", 64), ("torus", 32), ("random", 32)])
def test_eval_support_matches_oracle(ctx, name, R):
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    print(f"[eval] {c}", flush=True)
    g = _gpu(ctx, F)
    q = _queries(c, 20000, 1)
    out_val = 0.75
    fg = g.eval(q, c.end, _capi.EVAL_SUPPORT, out_val)
    fr = F.svo.eval_support(F.lam, q.astype(np.float64), c.end, out_val)
    err = np.abs(fg.astype(np.float64) - fr)
    print("max abs err", err.max(), "max |f|", np.abs(fr).max())
    assert err.max() <= TOL


@pytest.mark.parametrize("name,R", [("icosphere", 64), ("bunny", 64)])
def test_eval_dense_and_raw_match_bruteforce(ctx, name, R):
    """Dense semantics checked against the O(N n) brute-force sum itself."""
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    g = _gpu(ctx, F)
    rng = np.random.default_rng(5)
    q = rng.uniform(c.origin - 0.6 * c.width, c.end + 0.6 * c.width, size=(400, 3))
    fr = F.svo.eval_dense(F.lam, q)
    fg = g.eval(q, c.end, _capi.EVAL_RAW)
    assert np.abs(fg - fr).max() <= TOL
    fd = g.eval(q, c.end, _capi.EVAL_DENSE, 3.0)
    inside = np.all((q >= c.origin - c.width) & (q <= c.end + c.width), axis=1)
    assert np.abs(fd[inside] - fr[inside]).max() <= TOL
    assert np.all(fd[~inside] == np.float32(3.0))


def test_eval_at_mesh_vertices(ctx):
    """f at the model vertices through the fine node (cpLatentBSplineValue)."""
    from its_amd import _capi
    F = fitted("icosphere", 64)
    c = F.case
    g = _gpu(ctx, F)
    fr, rc = F.svo.eval_latent(F.lam, c.V)
    assert rc == 0
    fg = g.eval(c.V, c.end, _capi.EVAL_SUPPORT, 0.0)
    assert np.abs(fg - fr).max() <= TOL


def test_partition_of_unity(ctx):
    """lambda = 1: f(x) counts the levels whose cell containing x has a vertex
    corner, weighted by the hats of the present corners; at a level-0
    node interior every level contributes 1, so f = tree depth."""
    from its_amd import _capi
    c = case("icosphere", 32)
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, 32)
    g.set_lambda(np.ones(g.num_verts))
    # points at the centres of occupied voxels: every level's cell is a node
    vox = g.voxels()
    x = np.array([O.lib().ora_morton_encode(0, 0, 0)])  # warm library
    w0 = c.width[0] / 32
    xs, ys, zs = [], [], []
    from its_amd._capi import lib  # noqa
    for m in vox[:500]:
        xx = yy = zz = 0
        for b in range(10):
            xx |= ((int(m) >> (3 * b)) & 1) << b
            yy |= ((int(m) >> (3 * b + 1)) & 1) << b
            zz |= ((int(m) >> (3 * b + 2)) & 1) << b
        xs.append(xx); ys.append(yy); zs.append(zz)
    P = c.origin + (np.stack([xs, ys, zs], 1) + 0.5) * w0
    f = g.eval(P, c.end, _capi.EVAL_SUPPORT)
    np.testing.assert_allclose(f, g.depth, atol=1e-5)


def test_eval_empty_and_device_path(ctx):
    from its_amd import _capi
    F = fitted("icosphere", 64)
    g = _gpu(ctx, F)
    c = F.case
    assert len(g.eval(np.zeros((0, 3), np.float32), c.end)) == 0


@pytest.fixture
def restore_tunables():
    from its_amd import _capi
    keys = (_capi.TUNE_DENSE_CELLS, _capi.TUNE_EVAL_ILP, _capi.TUNE_EVAL_QUEUE, _capi.TUNE_EVAL_BLOCKS_PER_CU)
    old = {k: _capi.tune_get(k) for k in keys}
    yield
    for k, v in old.items():
        _capi.tune_set(k, v)


@pytest.mark.parametrize("dense_cells", [8, 512, 32768])
def test_eval_schemes_agree(ctx, restore_tunables, dense_cells):
    """Every launch scheme (per-thread walk with 1, 2 or 4 queries per thread,
    or the per-wavefront queue) and every collapse level gives the oracle's f.
    Off the upper box faces the schemes do the same binary32 arithmetic in the
    same order, so they agree bit for bit."""
    from its_amd import _capi
    F = fitted("bunny", 64)
    c = F.case
    _capi.tune_set(_capi.TUNE_DENSE_CELLS, dense_cells)
    g = _gpu(ctx, F)
    rng = np.random.default_rng(11)
    q = rng.uniform(c.origin - 0.05, c.end + 0.05, size=(30011, 3)).astype(np.float32)
    # points on the upper faces and corners of the grid box (u = R)
    face = rng.uniform(c.origin, c.end, size=(64, 3)).astype(np.float32)
    face[:, 0] = np.float32(c.end[0])
    face[:8] = np.float32(c.end)
    q = np.concatenate([q, face])
    fr = F.svo.eval_dense(F.lam, q[:2000].astype(np.float64))
    fr_face = F.svo.eval_dense(F.lam, face.astype(np.float64))
    ref = None
    for queue, ilp in [(0, 1), (0, 2), (0, 4), (1, 1)]:
        _capi.tune_set(_capi.TUNE_EVAL_QUEUE, queue)
        _capi.tune_set(_capi.TUNE_EVAL_ILP, ilp)
        for bpc in (1, 32):
            _capi.tune_set(_capi.TUNE_EVAL_BLOCKS_PER_CU, bpc)
            f = g.eval(q, c.end, _capi.EVAL_RAW)
            print(f"[eval scheme] dense={dense_cells} queue={queue} ilp={ilp} bpc={bpc}", flush=True)
            assert np.abs(f[:2000] - fr).max() <= TOL
            assert np.abs(f[-64:] - fr_face).max() <= TOL
            if ref is None:
                ref = f
            else:
                assert np.array_equal(f[:-64], ref[:-64])
    g.close()


@pytest.mark.parametrize("name,R", [("bunny", 64), ("torus", 32)])
def test_eval_boundary_points(ctx, name, R):
    """Points on the box faces (open box in SUPPORT mode, closed in DENSE),
    one binary32 step inside and outside them, and lattice points of every
    level (cell corners, where the walk picks a cell by rounding): binary32
    evaluation against the oracle's binary64 support sum at the same points."""
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    g = _gpu(ctx, F)
    rng = np.random.default_rng(3)
    o32, e32 = c.origin.astype(np.float32), c.end.astype(np.float32)
    pts = []
    for a in range(3):   # faces: the bound itself and its binary32 neighbours
        for bound in (c.origin[a], c.end[a]):
            for v in (np.float32(bound), np.nextafter(np.float32(bound), np.float32(-np.inf)),
                      np.nextafter(np.float32(bound), np.float32(np.inf))):
                p = rng.uniform(o32 + 0.1, e32 - 0.1, size=(40, 3)).astype(np.float32)
                p[:, a] = v
                pts.append(p)
    w = c.width[0] / R
    for d in range(int(np.log2(R)) + 1):   # lattice points of level d
        k = rng.integers(0, R // (1 << d) + 1, size=(200, 3))
        pts.append((c.origin + k * w * (1 << d)).astype(np.float32))
    P = np.concatenate(pts)
    out_val = -2.5
    fg = g.eval(P, c.end, _capi.EVAL_SUPPORT, out_val)
    fr = F.svo.eval_support(F.lam, P.astype(np.float64), c.end, out_val)
    # the open-box decision is exact: out_val on both sides at the same points
    np.testing.assert_array_equal(fg == np.float32(out_val), fr == out_val)
    assert np.abs(fg.astype(np.float64) - fr).max() <= TOL
    # closed box: the face points themselves are inside
    fd = g.eval(P, c.end, _capi.EVAL_DENSE, 7.0)
    Pd = P.astype(np.float64)
    inside = np.all((Pd >= c.origin - c.width) & (Pd <= c.end + c.width), axis=1)
    assert np.all(fd[~inside] == np.float32(7.0)) and np.all(fd[inside] != np.float32(7.0))
    g.close()


def test_hbm_bandwidth_probe(ctx):
    """The streaming-copy probe that the roofline report divides by: a
    positive rate below the published 8 TB/s, and range errors for sizes and
    repetition counts outside the documented bounds."""
    from its_amd import _capi
    print("[eval] streaming copy of 256 MiB", flush=True)
    g = ctx.hbm_bandwidth(256 << 20, 5)
    assert 100.0 < g < 8000.0, g
    for nbytes, reps in [(1 << 10, 5), (256 << 20, 0), (256 << 20, 101)]:
        with pytest.raises(_capi.ItsError) as e:
            ctx.hbm_bandwidth(nbytes, reps)
        assert e.value.code == -5
