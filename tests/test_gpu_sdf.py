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
