"""Shell iso band: intersection points and the inner / outer iso values.

Reference: ThinShells::cpIntersectionPoints (ThinShells.cpp:38-479) and the
min / max of cpLatentBSplineValue (655-755).  The oracle restates the
reference's loops (ordered edge set, voxel ranges, per-axis node-edge sweeps
over end-sorted edges); the GPU enumerates candidates per edge and per
triangle instead, so the two are compared as point sets.
"""
import numpy as np
import pytest

import oracle as O
from helpers import case, fitted

CASES = [("icosphere", 32), ("bunny", 32), ("torus", 16), ("knot_small", 32)]


def _sorted(P):
    return P[np.lexsort((P[:, 2], P[:, 1], P[:, 0]))]


@pytest.mark.parametrize("name,R", CASES[:2])
def test_oracle_points_on_mesh_and_lattice(name, R):
    c = case(name, R)
    s = O.SVO(c.tris, c.origin, c.width, R)
    P, ne = O.intersections(s, c.V, c.F)
    assert ne > 0 and len(P) > ne
    # every point is on the surface
    assert np.abs(O.signed_distance(c.V, c.F, P)).max() < 1e-12
    # edge points lie on a level-0 face plane, face points on a level-0 edge line
    w = c.width[0] / R
    u = (P - c.origin) / w
    on = np.abs(u - np.rint(u)) < 1e-6
    assert np.all(on[:ne].sum(1) >= 1)
    assert np.all(on[ne:].sum(1) >= 2)


@pytest.mark.gpu
@pytest.mark.parametrize("name,R", CASES)
def test_gpu_points_match_oracle(ctx, name, R):
    from its_amd import _capi
    c = case(name, R)
    print(f"[shell] {c}", flush=True)
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    s = O.SVO(c.tris, c.origin, c.width, R)
    P, ne = g.shell_points(c.V, c.F)
    Q, qe = O.intersections(s, c.V, c.F)
    # edge points: the GPU keeps duplicates, the reference removes exact ones
    Pe = np.unique(P[:ne], axis=0)
    Qe = np.unique(Q[:qe], axis=0)
    assert len(Pe) == len(Qe)
    assert np.array_equal(Pe, Qe)
    # face points: same multiset; end points of equal node edges may come from
    # different nodes, a difference of an ulp
    assert len(P) - ne == len(Q) - qe
    d = np.abs(_sorted(P[ne:]) - _sorted(Q[qe:]))
    assert d.max() <= 1e-12 if len(d) else True
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name,R", [("icosphere", 32), ("bunny", 64)])
def test_gpu_iso_values(ctx, name, R):
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    g.set_lambda(F.lam)
    inner, outer, n = g.shell_iso(c.V, c.F)
    Q, _ = O.intersections(F.svo, c.V, c.F)
    pts = np.concatenate([c.V, Q])
    f, rc = F.svo.eval_latent(F.lam, pts)
    assert rc == 0
    assert n == len(Q)
    assert abs(inner - f.min()) <= 1e-5 and abs(outer - f.max()) <= 1e-5
    assert inner < 0 < outer
    g.close()
