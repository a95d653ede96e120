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
