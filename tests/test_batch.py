"""Many meshes per launch (GPU): per-mesh results equal the single-mesh path.

No reference counterpart (the reference fits one model per process).  The
one-launch fit runs, per mesh, the algorithm and stopping rule of
its_lscg_solve; only the summation order of its binary64 reductions differs,
so lambda agrees to the solver tolerance and f to 1e-5.  Meshes above the
one-launch cap go through the many-mesh solver, whose per-mesh blocks,
reductions and arithmetic are those of its_lscg_solve: their coefficients and
iteration counts are bit-identical.  The batched evaluation uses the
arithmetic of the per-thread walk, so f is bit-identical to
its_bspline_eval_device with one query per thread.
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
    """cap: meshes above it take the many-mesh solver beside the one-launch kernel."""
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
    torch.cuda.synchronize()   # the fill runs on torch's stream, the evaluation on the context's
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
        if g.num_verts > cap:   # the many-mesh solver
            assert np.array_equal(lam_b, lam_s), c
            for k in ("iterations", "refinements", "normal_rel", "residual_abs", "initial_normal_rel"):
                assert st_b[k] == st_s[k], (c, k, st_b[k], st_s[k])
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
    with pytest.raises(_capi.ItsError):
        batch.add(items[0][1])                   # the same octree twice would race in the fit
    st0, _ = batch.fit(tol=1e-6)
    st1, _ = batch.fit(tol=1e-6, warm=True)      # from the converged coefficients
    for a, b in zip(st0, st1):
        assert b["converged"] and b["iterations"] <= 1
    with pytest.raises(_capi.ItsError):
        batch.eval_device(0, 0, 0, np.zeros(len(items) + 1, np.int64), 0)   # fields not rebuilt yet
    batch.close()


def test_batch_from_thread_contexts(ctx):
    """Octrees built by two host threads, one context each, join a batch on
    the test's context: fit and evaluation equal those of octrees built on
    the batch's own context, bit for bit."""
    import threading

    import torch
    from its_amd import _capi
    ctxs = [_capi.Context(0) for _ in range(2)]
    built, errors = [None] * len(MESHES), []

    def worker(t):
        try:
            for m in range(t, len(MESHES), 2):
                name, R = MESHES[m]
                c = case(name, R)
                g = _capi.Octree(ctxs[t], c.tris, c.origin, c.width, R)
                built[m] = (c, g, g.sdf_at_vertices(c.V, c.F))
        except Exception as e:   # re-raised below
            errors.append(e)

    print("[batch] building on two host threads", flush=True)
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    if errors:
        raise errors[0]
    rng = np.random.default_rng(11)
    counts = [2000] * len(MESHES)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    Q = np.concatenate([rng.uniform(case(*k).origin - 0.05, case(*k).end + 0.05, size=(n, 3))
                        for k, n in zip(MESHES, counts)]).astype(np.float32)
    dq = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()

    _, busy, busy_b = _build(ctx)[1]
    busy.lscg(busy_b, tol=1e-6)
    nb = 4_000_000
    qb = torch.rand((3, nb), device="cuda")
    fbusy = torch.empty(nb, device="cuda")
    torch.cuda.synchronize()

    def fit_eval(items):
        batch = _capi.Batch(ctx)
        for _, g, _ in items:
            batch.add(g)
        # queued, unsynchronised work on the batch context's stream when the fit starts
        for _ in range(4):
            busy.eval_device(qb[0].data_ptr(), qb[1].data_ptr(), qb[2].data_ptr(), nb, busy.origin + busy.width,
                             fbusy.data_ptr())
        st, _ = batch.fit(tol=1e-6)
        for _ in range(4):
            busy.eval_device(qb[0].data_ptr(), qb[1].data_ptr(), qb[2].data_ptr(), nb, busy.origin + busy.width,
                             fbusy.data_ptr())
        batch.fields()
        df = torch.zeros(offs[-1], device="cuda")
        torch.cuda.synchronize()   # the fill runs on torch's stream, the evaluation on the context's
        batch.eval_device(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, df.data_ptr(), 0.5)
        lams = [g.get_lambda() for _, g, _ in items]
        batch.close()
        return st, lams, df.cpu().numpy()

    st_t, lam_t, f_t = fit_eval(built)
    st_r, lam_r, f_r = fit_eval(_build(ctx))
    for a, b in zip(built, _build(ctx)):
        np.testing.assert_array_equal(a[2], b[2])             # right-hand sides
    for s1, s2 in zip(st_t, st_r):
        assert s1["converged"] and s1["iterations"] == s2["iterations"]
    for a, b in zip(lam_t, lam_r):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(f_t, f_r)
    for _, g, _ in built:
        g.close()
    for c_ in ctxs:
        c_.close()


def test_many_mesh_warm_and_iteration_limit(ctx):
    """With the one-launch cap at 1 every mesh goes through the many-mesh
    solver.  Stopped by an iteration limit, and then restarted warm from
    those coefficients, its coefficients and statistics equal the
    single-mesh solver's bit for bit."""
    from its_amd import _capi
    items = _build(ctx)[:5]
    old_cap = _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS)
    _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, 1)
    keys = ("iterations", "refinements", "normal_rel", "residual_abs", "initial_normal_rel", "converged")
    try:
        single = [g.lscg(b, tol=1e-6, max_iter=5) for _, g, b in items]
        lam0 = [lam.copy() for lam, _ in single]
        warm = [g.lscg(b, x0=lam0[m], tol=1e-6) for m, (_, g, b) in enumerate(items)]
        batch = _capi.Batch(ctx)
        for _, g, _ in items:
            batch.add(g)
        st, _ = batch.fit(tol=1e-6, max_iter=5)
        for m, (c, g, _) in enumerate(items):
            assert np.array_equal(g.get_lambda(), single[m][0]), c
            assert all(st[m][k] == single[m][1][k] for k in keys), (c, st[m], single[m][1])
            assert st[m]["iterations"] <= 5
        for m, (_, g, _) in enumerate(items):
            g.set_lambda(lam0[m])
        st_w, _ = batch.fit(tol=1e-6, warm=True)
        for m, (c, g, _) in enumerate(items):
            assert np.array_equal(g.get_lambda(), warm[m][0]), c
            assert all(st_w[m][k] == warm[m][1][k] for k in keys), (c, st_w[m], warm[m][1])
            assert st_w[m]["converged"]
        batch.close()
    finally:
        _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, old_cap)
