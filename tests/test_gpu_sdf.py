"""Signed distance at octree vertices (right-hand side of the fit), GPU
against the oracle (replaces igl::signed_distance, ThinShells.cpp:481-509)."""
import numpy as np
import pytest

import oracle as O
from helpers import case, exact_winding_number, fitted

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


@pytest.mark.parametrize("name", ["icosphere", "torus", "knot", "flipped_sphere"])
def test_sdf_ray_signs(ctx, name):
    """Closed meshes: signs settled by the +z ray crossings (exact
    orientation tests, ITS_TUNE_SDF_RAY = 1) against the winding-number fan
    sum (ITS_TUNE_SDF_RAY = 0).  Away from the surface (|d| > 1e-9) the
    signed distances are identical bit for bit.  Within 1e-9 of it (here
    edge midpoints, rounded off their edges by about 1e-17) the binary32 fan
    sum can miss the side; wherever the two differ the ray's sign is checked
    against the exact winding number (rational arithmetic along an axis ray,
    tests/helpers.py), negative where it is at least 1."""
    from its_amd import _capi, meshes
    if name == "icosphere":
        V, F = meshes.noisy_icosphere(seed=3, level=4)
    elif name == "torus":
        V, F = meshes.torus(48, 96)
    elif name == "knot":
        V, F = meshes.torus_knot_tube(n_seg=400, n_ring=24)
        V = meshes.jitter(V, 1e-3, seed=7)
    else:   # inward orientation: winding number -1 inside, so no point is negative
        V, F = meshes.icosphere(3)
        F = F[:, ::-1].copy()
    V = _capi.model_to_unit_cube(V)
    rng = np.random.default_rng(11)
    g = np.linspace(-0.05, 1.05, 45)
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    P = np.concatenate([grid, rng.uniform(-0.2, 1.2, size=(20000, 3)), V,
                        0.5 * (V[F[:, 0]] + V[F[:, 1]]), V + rng.normal(scale=1e-6, size=V.shape)])
    old = _capi.tune_get(_capi.TUNE_SDF_RAY)
    try:
        _capi.tune_set(_capi.TUNE_SDF_RAY, 0)
        d_fan = ctx.signed_distance(V, F, P)
        _capi.tune_set(_capi.TUNE_SDF_RAY, 1)
        d_ray = ctx.signed_distance(V, F, P)
    finally:
        _capi.tune_set(_capi.TUNE_SDF_RAY, old)
    assert np.array_equal(np.abs(d_ray), np.abs(d_fan))
    # the two paths may differ only within 1e-9 of the surface, and there the
    # ray's side is the oracle's
    diff = np.nonzero(d_ray.view(np.uint64) != d_fan.view(np.uint64))[0]
    assert np.all(np.abs(d_fan[diff]) <= 1e-9)
    sel = diff if len(diff) <= 300 else rng.choice(diff, 300, replace=False)
    w = []
    for i in sel:
        wi = None
        for axis, sign in ((0, 1), (1, 1), (2, 1), (0, -1), (1, -1), (2, -1)):
            wi = exact_winding_number(V, F, P[i], axis, sign)
            if wi is not None:
                break
        w.append(wi)
    print(f"[sdf ray] {name}: {len(F)} triangles, {len(P)} points, {int((d_ray < 0).sum())} inside; "
          f"the fan sum's sign differs at {len(diff)}, all within 1e-9 of the surface", flush=True)
    assert all(wi is not None for wi in w)
    assert np.array_equal(np.signbit(d_ray[sel]), np.array(w) >= 1)
    if name == "flipped_sphere":
        # (within rounding of an edge line the ray leaves a point to the fan
        # sum, whose binary32 side is unreliable that close)
        assert not np.signbit(d_ray[np.abs(d_ray) > 1e-9]).any()
    else:
        assert (d_ray < 0).sum() > 1000
