"""The rational-arithmetic winding number that referees the signed-distance
ray test (tests/helpers.py), on the CPU: every axis ray, in both
directions, gives 1 inside and 0 outside a closed outward sphere and -1
inside the same sphere with its triangles reversed; the points of a nested
pair of spheres give 2 inside both.  A crossing sign or direction slip in the
helper fails one of these."""
import numpy as np

from helpers import exact_winding_number

RAYS = ((0, 1), (1, 1), (2, 1), (0, -1), (1, -1), (2, -1))


def _winding(V, F, p):
    return {exact_winding_number(V, F, p, a, s) for a, s in RAYS} - {None}


def test_sphere_inside_outside_and_flipped():
    from its_amd import meshes
    V, F = meshes.icosphere(2)
    rng = np.random.default_rng(5)
    inside = rng.uniform(-0.5, 0.5, size=(6, 3)) * 0.8
    outside = np.array([[1.6, 0.3, -0.2], [0.1, -1.4, 0.2], [0.05, 0.1, 2.0], [-1.3, -0.9, 0.4]])
    for p in inside:
        assert _winding(V, F, p) == {1}, p
        assert _winding(V, F[:, ::-1], p) == {-1}, p
    for p in outside:
        assert _winding(V, F, p) == {0}, p


def test_nested_spheres():
    from its_amd import meshes
    V, F = meshes.icosphere(2)
    V2 = np.concatenate([V, 0.5 * V])
    F2 = np.concatenate([F, F + len(V)])
    assert _winding(V2, F2, np.array([0.03, -0.02, 0.01])) == {2}
    assert _winding(V2, F2, np.array([0.0, 0.75, 0.05])) == {1}
