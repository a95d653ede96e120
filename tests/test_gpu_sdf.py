"""Signed distance at octree vertices (right-hand side of the fit), GPU
against the oracle (replaces igl::signed_distance, ThinShells.cpp:481-509)."""
import numpy as np
import pytest

import oracle as O
from helpers import case, fitted

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,R", [("icosphere", 64), ("bunny", 32), ("torus", 32)])
def test_sdf_at_vertices(ctx, name, R):
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    b = g.sdf_at_vertices(c.V, c.F)
    np.testing.assert_array_equal(np.sign(b), np.sign(F.b))
    assert np.abs(b - F.b).max() <= 1e-12


def test_sdf_sphere_closed_form(ctx):
    from its_amd import meshes
    V, Fc = meshes.icosphere(4)
    rng = np.random.default_rng(2)
    P = rng.uniform(-1.5, 1.5, size=(3000, 3))
    d = ctx.signed_distance(V, Fc, P)
    r = np.linalg.norm(P, axis=1)
    # inscribed polyhedron: within the chord sagitta of the unit sphere
    assert np.all(np.abs(d - (r - 1)) < 5e-3)
    far = np.abs(r - 1) > 0.01
    assert np.all(np.sign(d[far]) == np.sign(r[far] - 1))


def test_sdf_culling_keeps_exact_results(ctx):
    """Points on vertices and edge midpoints (ties between triangles), just off
    the surface and far away: the bounding-sphere culling must leave the
    binary64 distance identical to the full search of the oracle."""
    from its_amd import _capi
    c = case("bunny", 32)
    rng = np.random.default_rng(7)
    V, F = c.V, c.F
    mids = 0.5 * (V[F[:, 0]] + V[F[:, 1]])
    near = V[rng.choice(len(V), 500)] + 1e-4 * rng.standard_normal((500, 3))
    far = rng.uniform(-3, 4, size=(500, 3))
    P = np.concatenate([V[:500], mids[:500], near, far])
    d = ctx.signed_distance(V, F, P)
    ref = O.signed_distance(V, F, P)
    # the culling decides which triangles get the binary64 distance: exact
    assert np.abs(np.abs(d) - np.abs(ref)).max() <= 1e-12
    # the sign comes from the binary32 winding number, unchanged by the
    # culling; on the surface and where the open mesh leaves w near 1/2 it is
    # a tie, so compare it only where the binary64 winding number is clear
    w = np.array([O.winding_number(V, F, p) for p in P])
    clear = (np.abs(ref) > 1e-6) & (np.abs(w - 0.5) > 0.05)
    assert clear.sum() > 900
    np.testing.assert_array_equal(np.sign(d[clear]), np.sign(ref[clear]))


def _tree_cases():
    from its_amd import meshes
    V, F = meshes.icosphere(3)
    F = F.astype(np.int32)
    out = {}
    # unwelded triangle soup: every edge is open, no boundary fan is kept
    out["soup"] = (V[F].reshape(-1, 3), np.arange(3 * len(F), dtype=np.int32).reshape(-1, 3))
    # half the faces reversed: winding numbers between 0 and 1
    Ff = F.copy()
    Ff[::2] = Ff[::2, ::-1]
    out["flipped"] = (V, Ff)
    # a closed sphere with every face twice (winding 2 inside), plus
    # degenerate triangles (a repeated vertex, three collinear points)
    Fd = np.concatenate([F, F, [[0, 0, 1], [2, 2, 2]]]).astype(np.int32)
    out["doubled_degenerate"] = (V, Fd)
    # non-manifold: a second sphere sharing one edge, plus a fin of three
    # triangles on that edge
    V2 = V + np.array([2.0, 0, 0])
    a, b = F[0, 0], F[0, 1]
    Vn = np.concatenate([V, V2, [[0.5, 3.0, 0.0], [0.5, -3.0, 0.0], [0.5, 0.0, 3.0]]])
    F2 = F + len(V)
    F2[F2 == len(V) + a] = a
    F2[F2 == len(V) + b] = b
    m = 2 * len(V)
    out["nonmanifold"] = (Vn, np.concatenate([F, F2, [[a, b, m], [b, a, m + 1], [a, b, m + 2]]]).astype(np.int32))
    out["single"] = (np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]]), np.array([[0, 1, 2]], dtype=np.int32))
    return out


@pytest.mark.parametrize("name", ["soup", "flipped", "doubled_degenerate", "nonmanifold", "single"])
def test_sdf_tree_mesh_kinds(ctx, name):
    """The triangle tree's pruned nearest search and boundary-fan winding
    number against the brute-force oracle on meshes that are open, doubly
    covered, non-manifold, inconsistently oriented or degenerate."""
    V, F = _tree_cases()[name]
    print(f"[sdf tree] {name} V={len(V)} F={len(F)}", flush=True)
    rng = np.random.default_rng(11)
    lo, hi = V.min(0), V.max(0)
    P = np.concatenate([rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), size=(1500, 3)),
                        V[rng.choice(len(V), 300)] + 1e-3 * rng.standard_normal((300, 3)), V[:50]])
    d = ctx.signed_distance(V, F, P)
    ref = O.signed_distance(V, F, P)
    assert np.abs(np.abs(d) - np.abs(ref)).max() <= 1e-12
    w = np.array([O.winding_number(V, F, p) for p in P])
    clear = (np.abs(ref) > 1e-6) & (np.abs(w - 0.5) > 0.05)
    assert clear.sum() > 1000
    np.testing.assert_array_equal(np.sign(d[clear]), np.sign(ref[clear]))


def test_sdf_tree_large_mesh(ctx):
    """200k-triangle knot tube: deep tree, long fans."""
    from its_amd import meshes
    V, F = meshes.torus_knot_tube()
    F = F.astype(np.int32)
    print(f"[sdf tree] knot V={len(V)} F={len(F)}", flush=True)
    rng = np.random.default_rng(5)
    lo, hi = V.min(0), V.max(0)
    P = np.concatenate([rng.uniform(lo - 0.1, hi + 0.1, size=(400, 3)),
                        V[rng.choice(len(V), 200)] + 0.02 * rng.standard_normal((200, 3))])
    d = ctx.signed_distance(V, F, P)
    ref = O.signed_distance(V, F, P)
    assert np.abs(np.abs(d) - np.abs(ref)).max() <= 1e-12
    far = np.abs(ref) > 1e-6
    np.testing.assert_array_equal(np.sign(d[far]), np.sign(ref[far]))
