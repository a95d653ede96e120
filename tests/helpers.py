"""Shared inputs for the parity tests.

Every case is a seeded mesh, normalised like the reference constructor
ThinShells(file, R, R, R, is2UnitCube=true, 1.0): model2UnitCube followed by
setUniformBoundingBox.
"""
import functools
import os

import numpy as np

import oracle as O
from its_amd import meshes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Case:
    def __init__(self, name, V, F, R):
        self.name = name
        self.V = O.normalize(V)
        self.F = np.ascontiguousarray(F, np.int32)
        self.origin, self.end, self.width = O.uniform_bbox(self.V)
        self.tris = meshes.triangles(self.V, self.F)
        self.R = R

    def __repr__(self):
        return f"Case({self.name}, T={len(self.F)}, R={self.R})"


def random_triangles(seed, T, scale=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, size=(T, 1, 3))
    V = (c + scale * rng.standard_normal((T, 3, 3))).reshape(-1, 3)
    F = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    return V, F


def bunny():
    d = np.load(os.path.join(GOLDEN, "bunny.npz"))
    return d["V"], d["F"]


@functools.lru_cache(maxsize=None)
def case(name, R):
    if name == "icosphere":
        V, F = meshes.noisy_icosphere(seed=1)
    elif name == "icosphere2":
        V, F = meshes.noisy_icosphere(seed=2, level=3)
    elif name == "torus":
        V, F = meshes.torus(64, 32)
    elif name == "knot_small":
        V, F = meshes.torus_knot_tube(n_seg=400, n_ring=16)
    elif name == "knot":
        V, F = meshes.torus_knot_tube()
        V = O.normalize(V)
        V = meshes.jitter(V, 1e-3, seed=7)
    elif name == "random":
        V, F = random_triangles(3, 300)
    elif name == "bunny":
        V, F = bunny()
    elif name == "single":
        V = np.array([[0.1, 0.2, 0.3], [0.9, 0.25, 0.35], [0.4, 0.8, 0.6]])
        F = np.array([[0, 1, 2]], np.int32)
    elif name == "cube":
        # axis-aligned faces: exact-boundary voxel tests
        V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
        F = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7], [2, 7, 6],
                      [1, 2, 6], [1, 6, 5], [3, 0, 4], [3, 4, 7]], np.int32)
    elif name == "corner":
        # all geometry in one octant: a level below the root with one occupied cell
        V = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.05], [0.0, 0.2, 0.1], [3.0, 3.0, 3.0]])
        F = np.array([[0, 1, 2]], np.int32)
    else:
        raise KeyError(name)
    return Case(name, V, F, R)


class Fitted:
    """Oracle octree, right-hand side and a tightly converged FP64 fit."""

    def __init__(self, c):
        self.case = c
        self.svo = O.SVO(c.tris, c.origin, c.width, c.R)
        self.b = O.signed_distance(c.V, c.F, self.svo.vert_xyz)
        self.A = O.Matrix(self.svo)
        self.lam_tight, self.st_tight = self.A.lscg(self.b, tol=1e-12)
        self.lam, self.st = self.A.lscg(self.b, tol=1e-6)


@functools.lru_cache(maxsize=None)
def fitted(name, R):
    return Fitted(case(name, R))

REFERENCE_SRC = "/root/reference/src"
