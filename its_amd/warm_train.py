"""Train and evaluate the warm-start network on seeded noisy icospheres.

Workload (learned warm start): icospheres of level 4 (2,562 vertices, 5,120
triangles) with radial noise r = 1 + 0.05 sum_{l<=4} a_lm Y_lm, a_lm ~ N(0,1),
normalised to the unit cube; training fits at R = 128 and 256 (tree depth 7
and 8), held-out meshes at R = 256.

For every training mesh the GPU builds the octree, the signed-distance
right-hand side and a cold fit, maps the fit to its dual coefficients
(lambda = D A^T w), and samples rows of (features, w / cell width).  The
network is trained with the library's BF16 MFMA forward/backward kernels and
Adam.  Each held-out mesh is then solved cold and warm to the same tolerance;
the report gives iterations, wall times (prediction included in the warm
time) and the largest difference between the warm and cold fields.

    python -m its_amd.warm_train --train 2000 --test 200 --out its_amd/data/warm_mlp.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mesh(seed, ctx_lib):
    from its_amd import meshes
    V, F = meshes.noisy_icosphere(seed=seed)
    V = ctx_lib.model_to_unit_cube(V)
    origin, end, width = ctx_lib.uniform_bbox(V)
    return V, F, origin, end, width


def build(ctx, seed, R):
    from its_amd import _capi, meshes
    V, F, origin, end, width = mesh(seed, _capi)
    svo = _capi.Octree(ctx, meshes.triangles(V, F), origin, width, R)
    b = svo.sdf_at_vertices(V, F)
    return svo, V, F, b, origin, end


def training_rows(ctx, torch, seed, R, rows, tol, gen):
    svo, V, F, b, _, _ = build(ctx, seed, R)
    n = svo.num_verts
    lam, st = svo.lscg(b, tol=tol)
    w = svo.warm_dual(lam)
    X = torch.empty((n, 16), dtype=torch.int16, device="cuda")
    s = torch.empty(n, dtype=torch.float32, device="cuda")
    svo.warm_features(X.data_ptr(), s.data_ptr())
    t = torch.from_numpy(w).cuda().to(torch.float32) / s
    idx = torch.randint(0, n, (min(rows, n),), generator=gen, device="cuda")
    out = X[idx].clone(), t[idx].clone()
    svo.close()
    return out, n, st["iterations"]


def evaluate(ctx, torch, net, seed, R, tol, q):
    from its_amd import _capi
    svo, V, F, b, origin, end = build(ctx, seed, R)
    svo.lscg(b, tol=tol)                       # builds the solver structure once
    cold = [svo.lscg(b, tol=tol) for _ in range(3)]
    lam_c, st_c = cold[-1]
    pts = origin + q * (end - origin)
    f_c = svo.eval(pts.astype(np.float32), end, _capi.EVAL_SUPPORT)
    warm = [svo.lscg_mlp(net, tol=tol) for _ in range(3)]
    lam_w, st_w = warm[-1]
    f_w = svo.eval(pts.astype(np.float32), end, _capi.EVAL_SUPPORT)
    t_cold = float(np.median([s["seconds"] for _, s in cold]))
    t_warm = float(np.median([s["seconds"] + s["seconds_predict"] for _, s in warm]))
    t_pred = float(np.median([s["seconds_predict"] for _, s in warm]))
    r = dict(seed=seed, R=R, n=svo.num_verts, iters_cold=st_c["iterations"], iters_warm=st_w["iterations"],
             t_cold=t_cold, t_warm=t_warm, t_predict=t_pred, init_rel_warm=st_w["initial_normal_rel"],
             normal_rel_cold=st_c["normal_rel"], normal_rel_warm=st_w["normal_rel"],
             converged=bool(st_w["converged"] and st_c["converged"]), f_maxdiff=float(np.abs(f_w - f_c).max()))
    svo.close()
    return r


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--rows", type=int, default=4096, help="sampled rows per training mesh")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "data", "warm_mlp.npz"))
    ap.add_argument("--report", default=None, help="JSON report path")
    a = ap.parse_args(argv)

    import torch
    torch.cuda.init()
    from its_amd import _capi
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = _capi.Context(0)
    ctx.set_stream(stream.cuda_stream)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(a.seed)

    # ---- training set
    t0 = time.time()
    Xs, Ts, ns = [], [], []
    for k in range(a.train):
        R = 128 if k % 2 == 0 else 256
        (X, T), n, _ = training_rows(ctx, torch, 1000 + k, R, a.rows, a.tol, gen)
        Xs.append(X)
        Ts.append(T)
        ns.append(n)
        if (k + 1) % 100 == 0:
            log(f"[warm] training fits {k + 1}/{a.train}  ({time.time() - t0:.0f} s)")
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    del Xs, Ts
    N = len(T)
    log(f"[warm] {N} rows from {a.train} fits, target rms {float(T.pow(2).mean().sqrt()):.3f}")

    # ---- training
    net = _capi.Mlp(ctx, seed=a.seed)
    t1 = time.time()
    losses = []
    for step in range(a.steps):
        lr = a.lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * step / a.steps)))
        idx = torch.randint(0, N, (a.batch,), generator=gen, device="cuda")
        xb, tb = X[idx].contiguous(), T[idx].contiguous()
        loss = net.train_step(xb.data_ptr(), tb.data_ptr(), a.batch, lr)
        losses.append(loss)
        if (step + 1) % 1000 == 0:
            log(f"[warm] step {step + 1}/{a.steps} loss {np.mean(losses[-100:]):.4e}  ({time.time() - t1:.0f} s)")
    var = float(T.var())
    train_s = time.time() - t1
    params = net.get_params()
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.savez(a.out, params=params, train_fits=a.train, rows=N, steps=a.steps, final_loss=float(np.mean(losses[-100:])),
             target_var=var)
    log(f"[warm] saved {a.out}; relative loss {np.mean(losses[-100:]) / var:.3f}")

    # ---- held-out evaluation
    q = np.random.default_rng(12345).uniform(0, 1, size=(20000, 3))
    res = []
    for k in range(a.test):
        res.append(evaluate(ctx, torch, net, 500000 + k, 256, a.tol, q))
        if (k + 1) % 20 == 0:
            r = res[-1]
            log(f"[warm] test {k + 1}/{a.test}: iters {r['iters_cold']} -> {r['iters_warm']}, "
                f"{r['t_cold'] * 1e3:.2f} -> {r['t_warm'] * 1e3:.2f} ms")
    ic = np.array([r["iters_cold"] for r in res])
    iw = np.array([r["iters_warm"] for r in res])
    tc = np.array([r["t_cold"] for r in res])
    tw = np.array([r["t_warm"] for r in res])
    summary = dict(
        workload="noisy icosphere level 4, train R=128/256, test R=256", train_fits=a.train, test_meshes=a.test,
        rows=N, steps=a.steps, batch=a.batch, train_seconds=train_s, relative_train_loss=float(np.mean(losses[-100:]) / var),
        tol=a.tol, mean_iters_cold=float(ic.mean()), mean_iters_warm=float(iw.mean()),
        mean_ms_cold=float(tc.mean() * 1e3), mean_ms_warm=float(tw.mean() * 1e3),
        mean_ms_predict=float(np.mean([r["t_predict"] for r in res]) * 1e3),
        speedup_time=float(tc.sum() / tw.sum()), speedup_iters=float(ic.sum() / iw.sum()),
        all_converged=bool(all(r["converged"] for r in res)),
        max_f_diff_warm_vs_cold=float(max(r["f_maxdiff"] for r in res)),
        mean_n=float(np.mean([r["n"] for r in res])))
    print(json.dumps(summary))
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(dict(summary=summary, per_mesh=res), fh, indent=1)
    net.close()
    return summary


if __name__ == "__main__":
    main()
