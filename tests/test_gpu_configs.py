"""The batching and learned-warm-start workloads at (close to) their sizes.

- Batching: 32 seeded meshes (icospheres, tori, blob unions) with triangle
  counts log-uniform in [1k, 200k] and octree depth 5-8, generated as the
  256-mesh batch workload generates them.  Every mesh's batched fit and
  batched evaluation equal the single-mesh path (meshes above the one-launch
  size included).  No reference counterpart: the reference fits one model
  per process (ThinShells.cpp:615-617).
- Learned warm start: the trained network (its_amd/data/warm_mlp.npz) on 20
  held-out noisy icospheres at R = 256 (the held-out seeds of the training
  script), each fitted cold and warm to 1e-6; the warm fits reach the
  tolerance, keep the cold field, and the mean iteration count and wall time
  ratios are printed and recorded.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_workload_32_meshes(ctx):
    import torch

    from its_amd import _capi, batch_bench, meshes

    rng = np.random.default_rng(0)
    specs = [batch_bench.make_mesh(k, rng) for k in range(32)]
    items = []
    for k, (kind, V, F, depth) in enumerate(specs):
        V = _capi.model_to_unit_cube(V)
        origin, end, width = _capi.uniform_bbox(V)
        g = _capi.Octree(ctx, meshes.triangles(V, F), origin, width, 1 << depth)
        b = g.sdf_at_vertices(V, F)
        items.append(dict(kind=kind, T=len(F), depth=depth, g=g, b=b, origin=origin, end=end))
    cap = _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS)
    big = [it for it in items if it["g"].num_verts > cap]
    print(f"[batch32] depths {sorted(set(it['depth'] for it in items))}, triangles "
          f"{min(it['T'] for it in items)}..{max(it['T'] for it in items)}, {len(big)} meshes above {cap} unknowns",
          flush=True)
    assert {5, 8} <= set(it["depth"] for it in items) and len(big) >= 2
    single = [it["g"].lscg(it["b"], tol=1e-6) for it in items]
    batch = _capi.Batch(ctx)
    for it in items:
        batch.add(it["g"])
    stats, sec = batch.fit(tol=1e-6)
    batch.fields()
    print(f"[batch32] batched fit {sec * 1e3:.1f} ms", flush=True)
    nq = 4096
    offs = np.arange(len(items) + 1, dtype=np.int64) * nq
    Q = np.concatenate([rng.uniform(it["origin"] - 0.05, it["end"] + 0.05, size=(nq, 3)) for it in items])
    Q = Q.astype(np.float32)
    dq = torch.from_numpy(np.ascontiguousarray(Q.T)).cuda()
    df = torch.zeros(offs[-1], device="cuda")
    torch.cuda.synchronize()
    batch.eval_device(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, df.data_ptr(), 0.5)
    fb = df.cpu().numpy()
    old = (_capi.tune_get(_capi.TUNE_EVAL_QUEUE), _capi.tune_get(_capi.TUNE_EVAL_ILP))
    _capi.tune_set(_capi.TUNE_EVAL_QUEUE, 0)
    _capi.tune_set(_capi.TUNE_EVAL_ILP, 1)
    try:
        for m, it in enumerate(items):
            assert stats[m]["converged"] and stats[m]["normal_rel"] <= 1e-6
            lam_b = it["g"].get_lambda()
            lam_s, st_s = single[m]
            assert np.abs(lam_b - lam_s).max() <= 1e-5 * max(np.abs(lam_s).max(), 1e-30), m
            if it["g"].num_verts > cap:   # many-mesh solver: bit-identical to the single-mesh one
                assert np.array_equal(lam_b, lam_s) and stats[m]["iterations"] == st_s["iterations"], m
            f_single = it["g"].eval(Q[offs[m]:offs[m + 1]], it["end"], _capi.EVAL_SUPPORT, 0.5)
            assert np.array_equal(fb[offs[m]:offs[m + 1]], f_single), m
    finally:
        _capi.tune_set(_capi.TUNE_EVAL_QUEUE, old[0])
        _capi.tune_set(_capi.TUNE_EVAL_ILP, old[1])
    batch.close()
    for it in items:
        it["g"].close()


def test_trained_warm_start_held_out(ctx):
    import torch

    from its_amd import _capi, warm_train

    wfile = os.path.join(ROOT, "its_amd", "data", "warm_mlp.npz")
    net = _capi.Mlp(ctx)
    net.set_params(np.load(wfile)["params"])
    untrained = _capi.Mlp(ctx, seed=5)
    q = np.random.default_rng(12345).uniform(0, 1, size=(20000, 3))
    rs, ru = [], []
    for k in range(20):
        rs.append(warm_train.evaluate(ctx, torch, net, 500000 + k, 256, 1e-6, q))
        ru.append(warm_train.evaluate(ctx, torch, untrained, 500000 + k, 256, 1e-6, q))
    ic = np.array([r["iters_cold"] for r in rs], float)
    iw = np.array([r["iters_warm"] for r in rs], float)
    iu = np.array([r["iters_warm"] for r in ru], float)
    tc = sum(r["t_cold"] for r in rs)
    tw = sum(r["t_warm"] for r in rs)
    print(f"[warm start] 20 held-out meshes at R=256: iterations cold {ic.mean():.2f}, trained {iw.mean():.2f}, "
          f"untrained {iu.mean():.2f}; wall time cold/warm {tc / tw:.3f} "
          f"(prediction {np.mean([r['t_predict'] for r in rs]) * 1e3:.3f} ms)", flush=True)
    assert all(r["converged"] for r in rs + ru)
    assert max(r["f_maxdiff"] for r in rs + ru) <= 1e-5
    # the trained network starts closer than the untrained one
    assert np.mean([r["init_rel_warm"] for r in rs]) < np.mean([r["init_rel_warm"] for r in ru])
    assert iw.mean() <= ic.mean()
    net.close()
    untrained.close()


def test_precision_blob_512_against_oracle(ctx):
    """The precision workload at full size: the blob union of 32 spheres
    (seed 42) remeshed to about 1.5 M triangles, R = 512 (ten levels).
    - octree: every node and vertex field bit for bit against the oracle's
      build (SparseVoxelOctree::createOctree, src/cuAcc/SVO.cu:674-1113);
    - fit: Eigen's stopping quantity ||A^T(b - A lambda)|| / ||A^T b||
      recomputed in binary64 with the oracle's assembled matrix
      (cpCoefficients, src/ThinShells.cpp:511-617) is below 1e-6;
    - evaluation: f at 20,000 octree vertices and 20,000 uniform points,
      binary32 on the GPU, against the oracle's binary64 support evaluation
      with the same coefficients (mt_cpuTest_2, src/ThinShells.cpp:1249-1294),
      within 1e-5."""
    import oracle as O
    import precision_report as P

    from its_amd import _capi, meshes

    R = 512
    (V, F), mres = P.blob_mesh(42, 1.5e6)
    V = _capi.model_to_unit_cube(V)
    origin, end, width = _capi.uniform_bbox(V)
    tris = meshes.triangles(V, F)
    print(f"[blob512] {len(F)} triangles (grid {mres}^3); GPU octree, fit and oracle octree", flush=True)
    assert 1.2e6 <= len(F) <= 1.8e6
    g = _capi.Octree(ctx, tris, origin, width, R)
    b = g.sdf_at_vertices(V, F)
    lam, st = g.lscg(b, tol=1e-6)
    ref = O.SVO(tris, origin, width, R)
    assert g.depth == ref.depth == 10
    assert g.num_nodes == ref.num_nodes and g.num_verts == ref.num_verts
    nd = g.nodes()
    for k, r in [("morton", ref.morton), ("is_leaf", ref.is_leaf), ("parent", ref.parent), ("childs", ref.childs),
                 ("neighbors", ref.neighbors)]:
        np.testing.assert_array_equal(nd[k], r, err_msg=k)
    np.testing.assert_array_equal(nd["origin"].view(np.uint64), ref.node_origin.view(np.uint64))
    np.testing.assert_array_equal(nd["width"].view(np.uint64), ref.node_width.view(np.uint64))
    xyz, owner = g.verts()
    np.testing.assert_array_equal(owner, ref.vert_owner)
    np.testing.assert_array_equal(xyz.view(np.uint64), ref.vert_xyz.view(np.uint64))
    print(f"[blob512] octree bit-exact: {g.num_nodes} nodes, {g.num_verts} vertices; assembling the oracle matrix",
          flush=True)
    A = O.Matrix(ref)
    r = b - A.mul(lam)
    g_rel = np.linalg.norm(A.tmul(r)) / np.linalg.norm(A.tmul(b))
    print(f"[blob512] fit: {st['iterations']} iterations, normal_rel {st['normal_rel']:.3e} (library), "
          f"{g_rel:.3e} (oracle matrix); residual_rel {np.linalg.norm(r) / np.linalg.norm(b):.3e}", flush=True)
    assert st["converged"] and g_rel < 1e-6
    del A
    rng = np.random.default_rng(5)
    sel = rng.choice(g.num_verts, 20_000, replace=False)
    pts = np.concatenate([xyz[sel], rng.uniform(origin - 0.05, end + 0.05, size=(20_000, 3))])
    p32 = pts.astype(np.float32)
    f = g.eval(p32, end, _capi.EVAL_SUPPORT, 0.0)
    f_ref = ref.eval_support(lam, p32.astype(np.float64), end, 0.0)
    err = np.abs(f - f_ref)
    print(f"[blob512] f at 20,000 vertices: max err {err[:20_000].max():.3e}; at 20,000 uniform points: "
          f"{err[20_000:].max():.3e}", flush=True)
    assert err.max() <= 1e-5
    g.close()
