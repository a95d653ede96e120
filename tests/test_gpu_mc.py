"""Marching cubes on the GPU against the oracle restatement of
MC::marching_cubes (MarchingCubes.cu:672-714)."""
import numpy as np
import pytest

import oracle as O
from helpers import fitted

pytestmark = pytest.mark.gpu


def _gpu(ctx, F):
    from its_amd import _capi
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)
    g.set_lambda(F.lam)
    return g


@pytest.mark.parametrize("res", [32, 48, 64])
def test_mc_matches_oracle(ctx, res):
    F = fitted("icosphere", 64)
    c = F.case
    print(f"[mc] {c} res={res}", flush=True)
    g = _gpu(ctx, F)
    R3 = (res, res, res)
    lat = g.mc_lattice(c.origin, c.width, R3)
    # lattice values against the oracle's dense evaluation at the same points
    ax = [c.origin[a] + np.arange(res + 1) * (c.width[a] / res) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    ref = F.svo.eval_dense_fast(F.lam, P)
    assert np.abs(lat - ref).max() <= 1e-5
    for iso in [0.0, 0.01]:
        soup_r, cube_r, nvalid_r = O.marching_cubes(F.svo, F.lam, c.origin, c.width, R3, iso)
        tri, st = g.marching_cubes(c.origin, c.width, R3, iso)
        # classification agrees except where a lattice value sits within
        # float rounding of the iso value
        near = np.abs(ref - iso) < 1e-5
        if not near.any():
            assert st["num_valid_voxels"] == nvalid_r
            assert st["num_verts"] == len(soup_r)
            err = np.abs(tri.astype(np.float64) - soup_r).max()
            print("iso", iso, "verts", len(tri), "max pos err", err)
            assert err <= 1e-4 * c.width[0]
        else:
            assert abs(st["num_verts"] - len(soup_r)) <= 12 * near.sum()


def test_mc_closed_surface(ctx):
    """A closed level set: every soup edge is shared by two triangles."""
    F = fitted("icosphere", 64)
    c = F.case
    g = _gpu(ctx, F)
    tri, st = g.marching_cubes(c.origin, c.width, (40, 40, 40), 0.0)
    assert st["num_verts"] > 0 and st["num_verts"] % 3 == 0
    key = np.round(tri.astype(np.float64) / 1e-6).astype(np.int64)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    T = inv.reshape(-1, 3)
    edges = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
    edges = edges[edges[:, 0] != edges[:, 1]]
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)


@pytest.mark.parametrize("name,R", [("icosphere", 64), ("torus", 32), ("corner", 16)])
def test_lattice_f64_matches_oracle(ctx, name, R):
    """The binary64 dense cascade on the octree's level-0 lattice (the
    (R+1)^3 points of the grid box) against the oracle's binary64 evaluation
    at the same points: different summation orders, same sums."""
    F = fitted(name, R)
    c = F.case
    print(f"[mc] lattice {c}", flush=True)
    g = _gpu(ctx, F)
    lat = g.lattice_f64()
    ax = [c.origin[a] + np.arange(R + 1) * (c.width[0] / R) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    ref = F.svo.eval_dense_fast(F.lam, P)
    assert np.abs(lat - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # the binary32 lattice of marching cubes is the same values, rounded once
    lat32 = g.mc_lattice(c.origin, c.width, (R, R, R))
    np.testing.assert_array_equal(lat32, lat.astype(np.float32))


@pytest.mark.parametrize("res", [(7, 13, 5), (65, 3, 20), (1, 1, 1)])
def test_mc_ragged_grids(ctx, res):
    """Grids whose rows do not fill whole waves, with uneven axes."""
    F = fitted("icosphere", 64)
    c = F.case
    print(f"[mc] ragged res={res}", flush=True)
    g = _gpu(ctx, F)
    iso = 0.0
    soup_r, cube_r, nvalid_r = O.marching_cubes(F.svo, F.lam, c.origin, c.width, res, iso)
    lat = g.mc_lattice(c.origin, c.width, res).astype(np.float64)
    near = np.abs(lat - iso) < 1e-5
    tri, st = g.marching_cubes(c.origin, c.width, res, iso)
    if not near.any():
        assert st["num_valid_voxels"] == nvalid_r and st["num_verts"] == len(soup_r)
        if len(soup_r):
            assert np.abs(tri.astype(np.float64) - soup_r).max() <= 1e-4 * c.width[0]


@pytest.mark.parametrize("res", [48, 64])
def test_mc_lattice_follows_coefficients(ctx, res):
    """The octree keeps its last marching-cubes lattice for the next call on
    the same grid; new coefficients must not reuse it (res 64 is the octree's
    own lattice, the dense cascade; res 48 goes through the field walk)."""
    F = fitted("icosphere", 64)
    c = F.case
    print(f"[mc] lattice cache res={res}", flush=True)
    g = _gpu(ctx, F)
    R3 = (res, res, res)
    ax = [c.origin[a] + np.arange(res + 1) * (c.width[a] / res) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    lat1 = g.mc_lattice(c.origin, c.width, R3)
    tri1, st1 = g.marching_cubes(c.origin, c.width, R3, 0.0)
    tri1b, _ = g.marching_cubes(c.origin, c.width, R3, 0.0)   # kept lattice
    np.testing.assert_array_equal(tri1, tri1b)
    lam2 = F.lam + 0.01
    g.set_lambda(lam2)
    lat2 = g.mc_lattice(c.origin, c.width, R3)
    assert np.abs(lat2 - F.svo.eval_dense_fast(lam2, P)).max() <= 1e-5
    assert np.abs(lat1 - F.svo.eval_dense_fast(F.lam, P)).max() <= 1e-5
    tri2, st2 = g.marching_cubes(c.origin, c.width, R3, 0.0)
    assert st2["num_verts"] != st1["num_verts"] or not np.array_equal(tri1, tri2)
    # a solve writes the coefficients too
    lam3, _ = g.lscg(F.b, tol=1e-6)
    lat3 = g.mc_lattice(c.origin, c.width, R3)
    assert np.abs(lat3 - F.svo.eval_dense_fast(lam3, P)).max() <= 1e-5


def test_mc_several_iso_values_and_coefficients(ctx):
    """its_mc_extract with an iso list and new coefficients (the three
    extractions of mcVisualization, ThinShells.cpp:1086-1129): the soups
    equal one extraction per value, in order."""
    F = fitted("icosphere", 64)
    c = F.case
    from its_amd import _capi
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)
    isos = [-0.01, 0.0, 0.015]
    soups, st = g.marching_cubes_multi(c.origin, c.width, (48, 48, 48), isos, lam=F.lam)
    assert st["num_verts"] == sum(len(s) for s in soups)
    for iso, soup in zip(isos, soups):
        one, _ = g.marching_cubes(c.origin, c.width, (48, 48, 48), iso)
        assert np.array_equal(one, soup)
    # the coefficients passed in are the octree's from then on
    np.testing.assert_array_equal(g.get_lambda(), F.lam)


@pytest.mark.parametrize("shuffle", [False, True])
def test_octree_from_vertices(ctx, shuffle):
    """An octree object from the vertex arrays MC::marching_cubes takes
    (nodeVertexArray, nodeWidthArray; MarchingCubes.h:136-139): levels from
    the owner widths, lattice keys from the positions.  Evaluation and the
    soup are bit-identical to the octree the arrays came from, in the
    reference's vertex order and in a shuffled one."""
    from its_amd import _capi
    F = fitted("icosphere", 64)
    c = F.case
    g = _gpu(ctx, F)
    xyz, owner = g.verts()
    widths = g.nodes()["width"]
    lam = F.lam
    if shuffle:
        perm = np.random.default_rng(5).permutation(len(xyz))
        xyz, owner, lam = xyz[perm], owner[perm], lam[perm]
    v = _capi.Octree.from_vertices(ctx, xyz, owner, widths)
    assert v.num_verts == g.num_verts and v.depth == g.depth
    np.testing.assert_array_equal(v.origin.view(np.uint64), c.origin.view(np.uint64))
    v.set_lambda(lam)
    np.testing.assert_array_equal(v.get_lambda(), lam)
    q = np.random.default_rng(2).uniform(c.origin - 0.05, c.end + 0.05, size=(20000, 3)).astype(np.float32)
    for mode in (_capi.EVAL_RAW, _capi.EVAL_DENSE, _capi.EVAL_SUPPORT):
        np.testing.assert_array_equal(v.eval(q, c.end, mode, 0.25), g.eval(q, c.end, mode, 0.25))
    a, _ = v.marching_cubes(c.origin, c.width, (64, 64, 64), 0.0)
    b, _ = g.marching_cubes(c.origin, c.width, (64, 64, 64), 0.0)
    assert len(a) > 0
    np.testing.assert_array_equal(a, b)
    with pytest.raises(_capi.ItsError):
        v.lscg(np.zeros(v.num_verts))          # no solver on a vertex-only octree
    v.close()


def test_octree_from_vertices_errors(ctx):
    from its_amd import _capi
    F = fitted("icosphere", 16)
    g = _gpu(ctx, F)
    xyz, owner = g.verts()
    widths = g.nodes()["width"].copy()
    bad = widths.copy()
    bad[owner[0]] *= 1.5                        # not the finest width times a power of two
    with pytest.raises(_capi.ItsError):
        _capi.Octree.from_vertices(ctx, xyz, owner, bad)
    dup = np.concatenate([xyz, xyz[:1]])        # a vertex twice on one level
    with pytest.raises(_capi.ItsError):
        _capi.Octree.from_vertices(ctx, dup, np.concatenate([owner, owner[:1]]), widths)
    off = xyz.copy()
    off[-1] += 0.3 * widths[owner[-1]]          # off its level's lattice
    with pytest.raises(_capi.ItsError):
        _capi.Octree.from_vertices(ctx, off, owner, widths)


def test_mc_soup_buffer_reuse(ctx):
    """The octree keeps its soup buffer between extractions and emits into it
    before the soup's size is known; a soup that does not fit is emitted
    again into a larger buffer.  Grids that alternate between small and large
    soups give, call after call, the soup of a fresh octree."""
    F = fitted("icosphere", 64)
    c = F.case
    fresh = {}
    for res in (16, 64, 40):
        g = _gpu(ctx, F)
        fresh[res] = g.marching_cubes(c.origin, c.width, (res, res, res), 0.0)[0]
        g.close()
    g = _gpu(ctx, F)
    for res in (16, 64, 16, 40, 64, 64, 16):
        tri, st = g.marching_cubes(c.origin, c.width, (res, res, res), 0.0)
        assert st["num_verts"] == len(fresh[res])
        assert np.array_equal(tri, fresh[res]), res
    g.close()
