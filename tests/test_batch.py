"""Many meshes per launch (GPU): per-mesh results equal the single-mesh path.

No reference counterpart (the reference fits one model per process).  The
batched fit runs, per mesh, the algorithm and stopping rule of
its_lscg_solve; only the summation order of its binary64 reductions differs,
so lambda agrees to the solver tolerance and f to 1e-5.  The batched
evaluation uses the arithmetic of the per-thread walk, so f is bit-identical
to its_bspline_eval_device with one query per thread.
"""
import numpy as np
import pytest

from helpers import case

pytestmark = pytest.mark.gpu

MESHES = [("icosphere", 32), ("bunny", 64), ("torus", 16), ("random", 32), ("knot_small", 64), ("single", 8),
          ("corner", 16), ("icosphere2", 32)]


def _build(ctx):
    from its_amd import _capi
    out = []
    for name, R in MESHES:
        c = case(name, R)
        g = _capi.Octree(ctx, c.tris, c.origin, c.width, R)
        b = g.sdf_at_vertices(c.V, c.F)
        out.append((c, g, b))
    return out


@pytest.mark.parametrize("cap", [1 << 30, 3000])
def test_batch_fit_and_eval_match_single(ctx, cap):
    """cap: meshes above it take the multi-workgroup solver beside the launch."""
    import torch
    from its_amd import _capi
    old_cap = _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS)
    _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, cap)
    items = _build(ctx)
    print("[batch] meshes", [f"{c.name}/{c.R} n={g.num_verts}" for c, g, _ in items], flush=True)
    single = []
    for c, g, b in items:
        lam, st = g.lscg(b, tol=1e-6)
        single.append((lam, st))
    batch = _capi.Batch(ctx)
    for _, g, _ in items:
        batch.add(g)
    stats, sec = batch.fit(tol=1e-6)
    stats2, sec2 = batch.fit(tol=1e-6)          # replays the captured graph
    print("batch fit s", sec, sec2)
    batch.fields()
    rng = np.random.default_rng(3)
    counts = [int(rng.integers(0, 5000)) for _ in items]
    counts[2] = 0                                # a mesh without queries
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    Q = np.zeros((offs[-1], 3), np.float32)
    for m, (c, g, _) in enumerate(items):
        Q[offs[m]:offs[m + 1]] = rng.uniform(c.origin - 0.05, c.end + 0.05, size=(counts[m], 3))
    dq = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
    df = torch.full((max(offs[-1], 1),), float("nan"), device="cuda")
    batch.eval_device(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, df.data_ptr(), 0.5)
    fb = df.cpu().numpy()
    _capi.tune_set(_capi.TUNE_EVAL_QUEUE, 0)
    _capi.tune_set(_capi.TUNE_EVAL_ILP, 1)
    for m, (c, g, b) in enumerate(items):
        lam_s, st_s = single[m]
        st_b = stats[m]
        assert st_b["converged"] and st_b["normal_rel"] <= 1e-6, (c, st_b)
        assert st_b["iterations"] == stats2[m]["iterations"]
        lam_b = g.get_lambda()
        scale = max(np.abs(lam_s).max(), 1e-30)
        assert np.abs(lam_b - lam_s).max() <= 1e-5 * scale, c
        if counts[m]:
            q = Q[offs[m]:offs[m + 1]]
            f_single = g.eval(q, c.origin + c.width, _capi.EVAL_SUPPORT, 0.5)
            assert np.array_equal(fb[offs[m]:offs[m + 1]], f_single), c
    batch.close()
    _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, old_cap)


def test_batch_warm_and_errors(ctx):
    from its_amd import _capi
    items = _build(ctx)[:3]
    batch = _capi.Batch(ctx)
    with pytest.raises(_capi.ItsError):
        batch.fit()                              # empty batch
    for _, g, _ in items:
        batch.add(g)
    st0, _ = batch.fit(tol=1e-6)
    st1, _ = batch.fit(tol=1e-6, warm=True)      # from the converged coefficients
    for a, b in zip(st0, st1):
        assert b["converged"] and b["iterations"] <= 1
    with pytest.raises(_capi.ItsError):
        batch.eval_device(0, 0, 0, np.zeros(len(items) + 1, np.int64), 0)   # fields not rebuilt yet
    batch.close()
