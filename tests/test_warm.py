"""Warm start in the row space of the collocation matrix (GPU).

No reference counterpart: the reference starts its LSCG from zero
(ThinShells.cpp:615-617).  Jacobi-preconditioned CGLS keeps its iterates in
x0 + range(D A^T); a start x0 = D A^T w0 therefore converges to the same
solution as the cold start, and the fitted field f does not depend on w0.
These tests pin that property, the dual map lambda -> w, and the features.
"""
import numpy as np
import pytest

import oracle as O
from helpers import case, fitted

pytestmark = pytest.mark.gpu


def _gpu(ctx, c):
    from its_amd import _capi
    return _capi.Octree(ctx, c.tris, c.origin, c.width, c.R)


@pytest.mark.parametrize("name,R", [("icosphere", 32), ("bunny", 32), ("torus", 16)])
def test_dual_roundtrip(ctx, name, R):
    """D A^T dual(lambda*) = lambda* for the converged binary64 fit, and the
    group sums of the dual solve the dual system A D A^T w = b."""
    F = fitted(name, R)
    c = F.case
    print(f"[warm dual] {c}", flush=True)
    g = _gpu(ctx, c)
    lam = F.lam_tight
    w = g.warm_dual(lam)
    lam0 = g.warm_apply(w)
    assert np.linalg.norm(lam0 - lam) <= 1e-8 * np.linalg.norm(lam)
    # A D A^T w = A lambda* = b (up to the fit's residual)
    r = F.A.mul(lam0) - F.b
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(F.b)
    g.close()


def test_features(ctx):
    import torch
    F = fitted("icosphere", 32)
    c = F.case
    g = _gpu(ctx, c)
    b = g.sdf_at_vertices(c.V, c.F)
    n = g.num_verts
    dx = torch.zeros((n, 16), dtype=torch.int16, device="cuda")
    ds = torch.zeros(n, dtype=torch.float32, device="cuda")
    g.warm_features(dx.data_ptr(), ds.data_ptr())
    torch.cuda.synchronize()
    import mlp_oracle as M
    X = M.bf16_value(dx.cpu().numpy().view(np.uint16).reshape(n, 16))
    s = ds.cpu().numpy().astype(np.float64)
    lev = np.zeros(n, int)
    for d in range(g.depth):
        lev[g.esum_verts()[d]:g.esum_verts()[d + 1]] = d
    w0 = c.width[0] / c.R
    np.testing.assert_allclose(s, w0 * 2.0 ** lev, rtol=1e-6)
    # feature 4 is the vertex's own signed distance over its cell width
    np.testing.assert_array_equal(X[:, 4], M.bf16_value(M.bf16_bits((b / (w0 * 2.0 ** lev)).astype(np.float32))))
    assert np.all(np.abs(X[:, :3]) <= 1.0)
    np.testing.assert_array_equal(X[:, 3], M.bf16_value(M.bf16_bits((lev / 8.0).astype(np.float32))))
    nn = np.linalg.norm(X[:, 12:15], axis=1)
    assert np.all(np.abs(nn - 1) < 1e-2)
    assert np.all(X[:, 15] == 0)
    # every corner distance is the distance of some vertex of the same level
    bl = {d: set(np.round(M.bf16_value(M.bf16_bits((b[lev == d] / (w0 * 2.0 ** d)).astype(np.float32))), 12))
          for d in range(g.depth)}
    for i in np.random.default_rng(0).choice(n, 200, replace=False):
        for k in range(8):
            assert round(X[i, 4 + k], 12) in bl[lev[i]]
    g.close()


@pytest.mark.parametrize("name,R", [("bunny", 64), ("icosphere", 64)])
def test_row_space_start_keeps_the_field(ctx, name, R):
    """A noisy dual start and an untrained network both reach the tolerance
    and give the cold start's field to 1e-5."""
    from its_amd import _capi
    c = case(name, R)
    print(f"[warm field] {c}", flush=True)
    g = _gpu(ctx, c)
    b = g.sdf_at_vertices(c.V, c.F)
    lam_c, st_c = g.lscg(b, tol=1e-6)
    rng = np.random.default_rng(1)
    q = rng.uniform(c.origin, c.end, size=(20000, 3)).astype(np.float32)
    f_c = g.eval(q, c.end, _capi.EVAL_SUPPORT)
    w = g.warm_dual(lam_c)
    w0 = w * (1 + 0.2 * rng.standard_normal(len(w)))
    lam_w, st_w = g.lscg(b, x0=g.warm_apply(w0), tol=1e-6)
    f_w = g.eval(q, c.end, _capi.EVAL_SUPPORT)
    print("cold", st_c["iterations"], "noisy dual start", st_w["iterations"], "init", st_w["initial_normal_rel"])
    assert st_w["converged"] and st_w["normal_rel"] <= 1e-6
    assert np.abs(f_w - f_c).max() <= 1e-5
    net = _capi.Mlp(ctx, seed=5)
    lam_m, st_m = g.lscg_mlp(net, tol=1e-6)
    f_m = g.eval(q, c.end, _capi.EVAL_SUPPORT)
    print("untrained network start", st_m["iterations"], "predict s", st_m["seconds_predict"])
    assert st_m["converged"] and st_m["normal_rel"] <= 1e-6
    assert np.abs(f_m - f_c).max() <= 1e-5
    net.close()
    g.close()
