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


def test_oracle_classify_rule():
    """The class rule on its boundaries: a value equal to the outer iso value
    is 1, equal to the inner one -1, NaN 0 (ThinShells.cpp:1412-1417)."""
    inner, outer = -0.25, 0.5
    f = np.array([-1.0, -0.25, -0.2499999, 0.0, 0.4999999, 0.5, 2.0, np.nan, np.inf, -np.inf])
    assert O.classify(f, inner, outer).tolist() == [-1, -1, 0, 0, 0, 1, 1, 0, 1, -1]
    # a band of zero width (inner == outer) sends every value to one side
    assert O.classify(np.array([0.1, 0.2, 0.3]), 0.2, 0.2).tolist() == [-1, 1, 1]


@pytest.mark.gpu
@pytest.mark.parametrize("name,R", [("icosphere", 32), ("bunny", 64), ("knot_small", 32)])
def test_gpu_classes_match_oracle(ctx, name, R):
    """multiPointQuery's classes (ThinShells.cpp:1183-1420): the library's
    iso values, binary32 support evaluation and its_classify against the
    oracle's iso values, binary64 support evaluation and class rule, with the
    same coefficients.  Classes may differ only where a value lies within the
    evaluation tolerance of an iso value."""
    from its_amd import _capi
    F = fitted(name, R)
    c = F.case
    g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
    g.set_lambda(F.lam)
    inner, outer, _ = g.shell_iso(c.V, c.F)
    inner_r, outer_r = O.shell_iso(F.svo, F.lam, c.V, c.F)
    assert abs(inner - inner_r) <= 1e-5 and abs(outer - outer_r) <= 1e-5
    rng = np.random.default_rng(17)
    Q = rng.uniform(c.origin - 0.05, c.end + 0.05, size=(200_000, 3))
    # points near the surface too, where the classes change
    Q = np.concatenate([Q, c.V + rng.normal(scale=0.5 * c.width[0] / R, size=c.V.shape)])
    fq = g.eval(Q.astype(np.float32), c.end, _capi.EVAL_SUPPORT, np.float32(outer))
    cls = _capi.classify(fq, inner, outer)
    f_r = F.svo.eval_support(F.lam, Q.astype(np.float32).astype(np.float64), c.end, outer_r)
    cls_r = O.classify(f_r, inner_r, outer_r)
    near = (np.abs(f_r - inner_r) <= 2e-5) | (np.abs(f_r - outer_r) <= 2e-5)
    bad = (cls != cls_r) & ~near
    print(f"[classes] {name} R={R}: {len(Q)} queries, classes -1/0/1 = "
          f"{[(int((cls_r == k).sum())) for k in (-1, 0, 1)]}, {int((cls != cls_r).sum())} differ, all near an iso value",
          flush=True)
    assert not bad.any()
    assert (cls_r == 0).sum() > 100 and (cls_r == -1).sum() > 100 and (cls_r == 1).sum() > 100
    g.close()
