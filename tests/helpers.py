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


# setUniformBoundingBox (BaseModel.cpp:24-51) works partly in binary32: the
# longest side is found with the C fmaxf on the sides rounded to float, a side
# is skipped only when that float equals it exactly, the padding delta is a
# float, halved in float; the final 1/11 margin is binary64.  The expected
# boxes below are derived by hand from those rules.
F32_09 = 15099494 / 2 ** 24          # binary32 nearest to 0.9 (0.9 * 2^24 = 15099494.4)


def bbox_known_answers():
    """[(V, origin, end, width)] of setUniformBoundingBox for three point sets."""
    cases = []
    # (a) sides 1, 1/2, 1/4, all exact in binary32: x skipped, y padded by 1/4, z by 3/8
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.5, 0.25]])
    o = np.array([0.0, -0.25, -0.375])
    e = np.array([1.0, 0.75, 0.625])
    cases.append((V, o, e))
    # (b) side 0.1 is not a binary32 number: delta = float(1 - 0.1) = float(0.9),
    #     halved in binary32; a binary64 delta would give 0.45
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 1.0]])
    o = np.array([0.0, 0.0 - F32_09 / 2, 0.0])
    e = np.array([1.0, 0.1 + F32_09 / 2, 1.0])
    cases.append((V, o, e))
    # (c) the longest side 1 + 2^-30 rounds to 1.0f, which differs from it, so
    #     that side gets delta = float(1 - (1 + 2^-30)) = -2^-30: it shrinks by
    #     2^-31 at each end; the side 1 is exact and skipped
    V = np.array([[0.0, 0.0, 0.0], [1.0 + 2.0 ** -30, 1.0, 0.5]])
    o = np.array([2.0 ** -31, 0.0, -0.25])
    e = np.array([1.0 + 2.0 ** -30 - 2.0 ** -31, 1.0, 0.75])
    cases.append((V, o, e))
    out = []
    for V, o, e in cases:
        eps = (e - o) / 11
        out.append((V, o - eps, e + eps, (e + eps) - (o - eps)))
    return out



def exact_winding_number(V, F, p, axis=0, sign=1):
    """Winding number of a closed, consistently oriented triangle mesh at p,
    exactly: the signed count of crossings of the ray from p along axis in
    direction sign (+1 or -1), every orientation evaluated in rational
    arithmetic (fractions) from the binary64 coordinates.  A crossing counts
    the sign of the triangle's normal along the ray.  None when the ray meets
    an edge or p lies on a triangle's plane where it crosses (then another
    ray decides)."""
    from fractions import Fraction as Q

    V = np.asarray(V, np.float64)
    u, v = (axis + 1) % 3, (axis + 2) % 3   # (u, v, axis) is a right-handed frame
    P = [Q(float(x)) for x in p]
    T = V[np.asarray(F)]
    # triangles whose (u, v) box holds p and that reach beyond p along the ray
    lo, hi = T.min(1), T.max(1)
    ahead = (hi[:, axis] >= p[axis]) if sign > 0 else (lo[:, axis] <= p[axis])
    cand = np.nonzero((lo[:, u] <= p[u]) & (hi[:, u] >= p[u]) & (lo[:, v] <= p[v]) & (hi[:, v] >= p[v]) &
                      ahead)[0]

    def o2(a, b):
        return (a[u] - P[u]) * (b[v] - P[v]) - (a[v] - P[v]) * (b[u] - P[u])

    w = 0
    for t in cand:
        a, b, c = ([Q(float(x)) for x in T[t][k]] for k in range(3))
        s = [o2(a, b), o2(b, c), o2(c, a)]
        if any(x > 0 for x in s) and any(x < 0 for x in s):
            continue
        if any(x == 0 for x in s):
            return None
        # s[0]: the sign of the normal's axis component; along the ray: sg
        sg = (1 if s[0] > 0 else -1) * (1 if sign > 0 else -1)
        ad = [a[k] - P[k] for k in range(3)]
        bd = [b[k] - P[k] for k in range(3)]
        cd = [c[k] - P[k] for k in range(3)]
        n = [bd[(k + 1) % 3] * cd[(k + 2) % 3] - bd[(k + 2) % 3] * cd[(k + 1) % 3] for k in range(3)]
        det = sum(ad[k] * n[k] for k in range(3))
        if det == 0:
            return None
        if (det > 0) == (sg > 0):   # the plane is crossed ahead of p
            w += sg
    return w
