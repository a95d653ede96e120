"""Synthetic inputs: every generator yields a closed, outward-oriented
surface (winding number 1 inside, 0 outside), so the signed-distance
right-hand side is negative inside as the reference's fast-winding-number
sign convention gives (ThinShells.cpp:481-509)."""
import math

import numpy as np
import pytest

import oracle as O
from its_amd import grid_mc, meshes


def _interior_exterior(name):
    if name == "icosphere":
        V, F = meshes.noisy_icosphere(seed=1, level=3)
        return V, F, V.mean(0), V.mean(0) + 5.0
    if name == "torus":
        V, F = meshes.torus(32, 16)
        return V, F, np.array([1.0, 0, 0]), np.array([0.0, 0, 0])
    if name == "knot":
        V, F = meshes.torus_knot_tube(n_seg=300, n_ring=12)
        t = 0.3
        c = np.array([(2 + math.cos(5 * t)) * math.cos(3 * t), (2 + math.cos(5 * t)) * math.sin(3 * t), math.sin(5 * t)])
        return V, F, c, np.array([0.0, 0, 3.0])
    if name == "blob":
        V, F = grid_mc.blob_union(5, res=48, device="cpu")
        C, R = grid_mc.blob_union_sdf(5)
        return V, F, C[0], np.array([2.0, 2.0, 2.0])
    raise KeyError(name)


@pytest.mark.parametrize("name", ["icosphere", "torus", "knot", "blob"])
def test_generators_are_closed_and_outward(name):
    V, F, inside, outside = _interior_exterior(name)
    assert abs(O.winding_number(V, F, inside) - 1.0) < 1e-6
    assert abs(O.winding_number(V, F, outside)) < 1e-6
    d = O.signed_distance(V, F, np.stack([inside, outside]))
    assert d[0] < 0 < d[1]
    # closed: every edge is shared by exactly two triangles
    E = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(E, axis=0, return_counts=True)
    assert np.all(cnt == 2)
