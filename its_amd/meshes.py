"""Seeded synthetic meshes for the benchmark workloads.

Every generator is deterministic for a given seed.  Random numbers come from
numpy's PCG64 generator (``numpy.random.default_rng``).

Meshes are returned as ``(V, F)``:
- ``V`` is float64 with shape (nV, 3).
- ``F`` is int32 with shape (nF, 3).

Use :func:`triangles` to expand them into the ``T x 9`` (p1, p2, p3) layout.
That is the layout the octree builder takes, and it matches the reference's
``Triangle<V3d>`` list (src/BaseModel.cpp:784-803).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "icosphere",
    "noisy_icosphere",
    "torus_knot_tube",
    "torus",
    "blob_union",
    "triangles",
    "jitter",
]


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array(
        [
            [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
            [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
            [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
        ],
        dtype=np.float64,
    )
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    return V, F


def icosphere(level: int):
    """Unit icosphere with ``level`` 4-to-1 subdivisions.

    Level 4 has 2,562 vertices and 5,120 triangles.
    """
    V, F = _icosahedron()
    for _ in range(level):
        verts = [v for v in V]
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        V = np.array(verts)
        F = np.array(nf, dtype=np.int64)
    return V, F.astype(np.int32)


def _real_sph_harm(l: int, m: int, theta, phi):
    """Real spherical harmonic Y_lm for colatitude theta and azimuth phi.

    The normalisation is orthonormal on the sphere.
    """
    from scipy.special import lpmv

    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    P = lpmv(am, l, np.cos(theta))
    if m == 0:
        return norm * P
    if m > 0:
        return math.sqrt(2.0) * norm * P * np.cos(m * phi)
    return math.sqrt(2.0) * norm * P * np.sin(am * phi)


def noisy_icosphere(seed: int, level: int = 4, amp: float = 0.05, lmax: int = 4):
    """Icosphere with seeded radial noise.

    The radius is ``r = 1 + amp * sum_{l<=lmax, |m|<=l} a_lm Y_lm`` with
    ``a_lm ~ N(0, 1)``.  There are 25 real harmonics for ``lmax = 4``.
    """
    V, F = icosphere(level)
    rng = np.random.default_rng(seed)
    theta = np.arccos(np.clip(V[:, 2], -1.0, 1.0))
    phi = np.arctan2(V[:, 1], V[:, 0])
    r = np.ones(len(V))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            r += amp * rng.standard_normal() * _real_sph_harm(l, m, theta, phi)
    return V * r[:, None], F


def torus_knot_tube(n_seg: int = 2000, n_ring: int = 50, p: int = 3, q: int = 5, tube_r: float = 0.25):
    """Tube around the (p, q) torus knot.

    The centre curve is ``((2 + cos qt) cos pt, (2 + cos qt) sin pt, sin qt)``.
    The tube has ``2 * n_seg * n_ring`` triangles, which is 200,000 at the
    defaults.  The tube frame comes from Gram-Schmidt of the torus-radial
    direction against the curve tangent, so the frame is periodic and the tube
    closes without a seam.
    """
    t = np.arange(n_seg) * (2.0 * math.pi / n_seg)
    c = np.stack([(2 + np.cos(q * t)) * np.cos(p * t), (2 + np.cos(q * t)) * np.sin(p * t), np.sin(q * t)], 1)
    dc = np.stack(
        [
            -q * np.sin(q * t) * np.cos(p * t) - p * (2 + np.cos(q * t)) * np.sin(p * t),
            -q * np.sin(q * t) * np.sin(p * t) + p * (2 + np.cos(q * t)) * np.cos(p * t),
            q * np.cos(q * t),
        ],
        1,
    )
    T = dc / np.linalg.norm(dc, axis=1, keepdims=True)
    u = np.stack([np.cos(q * t) * np.cos(p * t), np.cos(q * t) * np.sin(p * t), np.sin(q * t)], 1)
    N = u - np.sum(u * T, 1, keepdims=True) * T
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    B = np.cross(T, N)
    a = np.arange(n_ring) * (2.0 * math.pi / n_ring)
    V = (c[:, None, :] + tube_r * (np.cos(a)[None, :, None] * N[:, None, :] + np.sin(a)[None, :, None] * B[:, None, :]))
    V = V.reshape(-1, 3)
    i = np.arange(n_seg)[:, None]
    j = np.arange(n_ring)[None, :]
    v00 = i * n_ring + j
    v10 = ((i + 1) % n_seg) * n_ring + j
    v01 = i * n_ring + (j + 1) % n_ring
    v11 = ((i + 1) % n_seg) * n_ring + (j + 1) % n_ring
    # counter-clockwise seen from outside: outward normals, winding number 1 inside
    F = np.concatenate([np.stack([v00, v11, v10], -1).reshape(-1, 3), np.stack([v00, v01, v11], -1).reshape(-1, 3)])
    return V, F.astype(np.int32)


def torus(n_u: int, n_v: int, R: float = 1.0, r: float = 0.35):
    """Ring torus with ``2 * n_u * n_v`` triangles."""
    u = np.arange(n_u) * (2 * math.pi / n_u)
    v = np.arange(n_v) * (2 * math.pi / n_v)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    V = np.stack([(R + r * np.cos(vv)) * np.cos(uu), (R + r * np.cos(vv)) * np.sin(uu), r * np.sin(vv)], -1).reshape(-1, 3)
    i = np.arange(n_u)[:, None]
    j = np.arange(n_v)[None, :]
    v00 = i * n_v + j
    v10 = ((i + 1) % n_u) * n_v + j
    v01 = i * n_v + (j + 1) % n_v
    v11 = ((i + 1) % n_u) * n_v + (j + 1) % n_v
    F = np.concatenate([np.stack([v00, v10, v11], -1).reshape(-1, 3), np.stack([v00, v11, v01], -1).reshape(-1, 3)])
    return V, F.astype(np.int32)


def blob_union(seed: int, n_spheres: int = 32, res: int = 64):
    """Union of seeded spheres, meshed by marching cubes of the union SDF.

    The SDF is ``min_i(|x - c_i| - r_i)`` with ``c_i ~ U[0.2, 0.8]^3`` and
    ``r_i ~ U[0.05, 0.15]``.  It is sampled on a ``res^3`` cell grid over
    [0, 1]^3, and the zero level set is extracted with the project's
    marching-cubes table.  Equal vertices are welded, so the mesh is indexed.
    """
    from ._mc_numpy import marching_cubes_grid

    rng = np.random.default_rng(seed)
    C = rng.uniform(0.2, 0.8, size=(n_spheres, 3))
    Rr = rng.uniform(0.05, 0.15, size=n_spheres)
    g = (np.arange(res + 1) + 0.0) / res
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X, Y, Z], -1).reshape(-1, 3)
    sdf = np.full(len(P), np.inf)
    for c, r in zip(C, Rr):
        sdf = np.minimum(sdf, np.linalg.norm(P - c, axis=1) - r)
    return marching_cubes_grid(sdf.reshape(res + 1, res + 1, res + 1), (0.0, 0.0, 0.0), 1.0 / res)


def jitter(V, sigma: float, seed: int):
    """Add seeded Gaussian vertex jitter N(0, sigma^2) per coordinate."""
    rng = np.random.default_rng(seed)
    return V + sigma * rng.standard_normal(V.shape)


def triangles(V, F):
    """Expand an indexed mesh into the T x 9 float64 triangle array (p1, p2, p3)."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F)
    return np.ascontiguousarray(V[F].reshape(-1, 9))
