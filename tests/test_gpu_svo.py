"""Octree build on the GPU against the CPU restatement: bit for bit.

The reference is SparseVoxelOctree::createOctree (src/cuAcc/SVO.cu:674-1113).
These tests compare every field of every SVONode (SVO.h:15-32), the
per-level vertex arrays and owners, and the voxel set.
"""
import numpy as np
import pytest

import oracle as O
from helpers import case

pytestmark = pytest.mark.gpu

CASES = [("icosphere", 64), ("icosphere", 16), ("torus", 32), ("random", 64), ("bunny", 64), ("single", 8),
         ("cube", 16), ("corner", 16), ("knot_small", 128)]


def _pair(ctx, name, R):
    from its_amd import _capi
    c = case(name, R)
    print(f"[svo] {c}", flush=True)
    ref = O.SVO(c.tris, c.origin, c.width, R)
    gpu = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    return c, ref, gpu


@pytest.mark.parametrize("name,R", CASES)
def test_svo_bit_exact(ctx, name, R):
    c, ref, gpu = _pair(ctx, name, R)
    assert gpu.num_voxels == ref.num_voxels
    assert gpu.depth == ref.depth
    np.testing.assert_array_equal(gpu.depth_num_nodes(), ref.depth_num_nodes[: ref.depth])
    np.testing.assert_array_equal(gpu.esum_verts(), ref.esum_verts)
    assert gpu.num_nodes == ref.num_nodes and gpu.num_verts == ref.num_verts
    nd = gpu.nodes()
    np.testing.assert_array_equal(nd["morton"], ref.morton)
    np.testing.assert_array_equal(nd["is_leaf"], ref.is_leaf)
    np.testing.assert_array_equal(nd["parent"], ref.parent)
    np.testing.assert_array_equal(nd["childs"], ref.childs)
    np.testing.assert_array_equal(nd["neighbors"], ref.neighbors)
    # binary64 geometry: identical bits
    np.testing.assert_array_equal(nd["origin"].view(np.uint64), ref.node_origin.view(np.uint64))
    np.testing.assert_array_equal(nd["width"].view(np.uint64), ref.node_width.view(np.uint64))
    xyz, owner = gpu.verts()
    np.testing.assert_array_equal(owner, ref.vert_owner)
    np.testing.assert_array_equal(xyz.view(np.uint64), ref.vert_xyz.view(np.uint64))
    vox = gpu.voxels()
    np.testing.assert_array_equal(vox, O.voxelize(c.tris, c.origin, c.width, R))


def test_svo_structural_invariants(ctx):
    """Size-independent checks at the full depth-8 torus-knot size (200k triangles)."""
    from its_amd import _capi
    c = case("knot", 256)
    print(f"[svo] {c}", flush=True)
    gpu = _capi.Octree(ctx, c.tris, c.origin, c.width, 256)
    assert gpu.depth == 9
    nd = gpu.nodes()
    L = gpu.depth
    dn = gpu.depth_num_nodes()
    off = np.concatenate([[0], np.cumsum(dn)])
    # nodes of a level are sorted by Morton code, octets complete
    for d in range(L):
        m = nd["morton"][off[d]:off[d + 1]]
        assert np.all(np.diff(m.astype(np.int64)) > 0)
        if dn[d] > 1:
            assert dn[d] % 8 == 0
            assert np.all((m.reshape(-1, 8) & 7) == np.arange(8))
    # parent / child links are mutually consistent
    ch = nd["childs"]
    has = ch[:, 0] != 0xFFFFFFFF
    for i in np.nonzero(has)[0][:20000]:
        assert np.all(nd["parent"][ch[i]] == i)
    assert nd["parent"][-1] == 0xFFFFFFFF
    # neighbour symmetry: k and 26-k are inverse offsets
    nb = nd["neighbors"]
    idx = np.arange(gpu.num_nodes)
    assert np.all(nb[:, 13] == idx)
    for k in range(27):
        j = nb[:, k]
        ok = j != 0xFFFFFFFF
        assert np.all(nb[j[ok], 26 - k] == idx[ok])
    # every vertex owner has the vertex as a corner
    xyz, owner = gpu.verts()
    o = nd["origin"][owner]
    w = nd["width"][owner][:, None]
    off_v = (xyz - o) / w
    assert np.all(np.isclose(off_v, 0) | np.isclose(off_v, 1))
    # voxel count matches the oracle's voxeliser (cheap to check at full size)
    np.testing.assert_array_equal(gpu.voxels(), O.voxelize(c.tris, c.origin, c.width, 256))


@pytest.mark.parametrize("name,R", [("icosphere", 16), ("bunny", 64), ("torus", 32), ("single", 8)])
def test_fine_edges_match_oracle(ctx, name, R):
    """Level-0 node edges (fineNodeEdgeArray, SVO.cu:965-1110): order, owner
    and end-point bits equal to the oracle's restatement."""
    from its_amd import _capi
    c = case(name, R)
    print(f"[fine edges] {c}", flush=True)
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    s = O.SVO(c.tris, c.origin, c.width, R)
    ends, owner = g.fine_edges()
    ends_o, owner_o = O.fine_edges(s)
    assert len(owner) == len(owner_o) > 0
    np.testing.assert_array_equal(owner, owner_o)
    np.testing.assert_array_equal(ends, ends_o)
    g.close()


@pytest.mark.parametrize("name,R", [("bunny", 64), ("torus", 32)])
def test_vertex_lattice_keys(ctx, name, R):
    """Vertex lattice keys: level d vertices sit in the d-th range of the
    global order, ascending in (x, y, z) (the order of depthNodeVertexArray,
    SVO.cu:1064-1096), at origin + key * (voxel width * 2^d) up to rounding."""
    from its_amd import _capi
    c = case(name, R)
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    ijk, lev = g.vert_lattice()
    xyz, _ = g.verts()
    es = g.esum_verts()
    for d in range(g.depth):
        sl = slice(int(es[d]), int(es[d + 1]))
        assert np.all(lev[sl] == d)
        k = ijk[sl].astype(np.int64)
        key = (k[:, 0] * (R + 1) + k[:, 1]) * (R + 1) + k[:, 2]
        assert np.all(np.diff(key) > 0)
        w = c.width[0] / R * 2.0 ** d
        assert np.abs(c.origin + k * w - xyz[sl]).max() <= 1e-12
    g.close()
