"""Structure of the warm-start network's error, and what it costs in LSCG
iterations.

For held-out noisy icospheres at R = 256 (the test meshes of
its_amd/warm_train.py) the report compares, per mesh:
- the cold solve (lambda_0 = 0) and the trained network's start;
- the exact dual w* of the converged fit (tolerance 1e-10), for which the
  start lambda_0 = D A^T w* is the solution itself;
- starts w* + s e for the network's own error e = w_net - w* scaled by
  s = 1 ... 1e-3, and the same with e replaced by i.i.d. noise of the same
  norm per level;
- starts that keep the network's error on one octree level only, and starts
  that remove it on one level only.
For each start: LSCG iterations to the normal residual 1e-6 (Eigen's stopping
quantity) and the initial normal residual.  The network's relative error per
level is reported too.  The solves are the product path (its_lscg_solve from
lambda_0, warm start); no reference counterpart (the reference starts from
zero, ThinShells.cpp:615-617).

    python tests/warm_report.py [--meshes 8] [--out report.json]
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


SCALES = (1.0, 0.3, 0.1, 0.03, 0.01, 1e-3)


def study_mesh(ctx, torch, net, seed, R=256, tol=1e-6):
    from its_amd import warm_train
    svo, V, F, b, origin, end = warm_train.build(ctx, seed, R)
    n = svo.num_verts
    _, lev = svo.vert_lattice()
    lam_c, st_c = svo.lscg(b, tol=tol)
    lam_t, _ = svo.lscg(b, tol=1e-10)
    w_star = svo.warm_dual(lam_t)
    X = torch.empty((n, 16), dtype=torch.int16, device="cuda")
    s = torch.empty(n, dtype=torch.float32, device="cuda")
    svo.warm_features(X.data_ptr(), s.data_ptr())
    y = torch.empty(n, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    net.forward_device(X.data_ptr(), n, y.data_ptr())
    torch.cuda.synchronize()
    w_net = (y.double() * s.double()).cpu().numpy()
    e = w_net - w_star
    rng = np.random.default_rng(seed)
    xi = rng.standard_normal(n)
    levels = sorted(set(lev.tolist()))
    for k in levels:   # i.i.d. noise with the network error's norm on each level
        m = lev == k
        xi[m] *= np.linalg.norm(e[m]) / max(np.linalg.norm(xi[m]), 1e-300)

    def start(w):
        _, st = svo.lscg(b, x0=svo.warm_apply(w), tol=tol)
        return {"iterations": st["iterations"], "initial_normal_rel": st["initial_normal_rel"],
                "converged": bool(st["converged"])}

    rep = {"seed": seed, "n": n, "cold": {"iterations": st_c["iterations"], "initial_normal_rel": 1.0},
           "exact_dual": start(w_star), "network": start(w_net),
           "relative_error": float(np.linalg.norm(e) / np.linalg.norm(w_star)),
           "levels": {}, "scaled_network_error": {}, "scaled_iid_error": {}}
    for sc in SCALES:
        rep["scaled_network_error"][str(sc)] = start(w_star + sc * e)
        rep["scaled_iid_error"][str(sc)] = start(w_star + sc * xi)
    for k in levels:
        m = lev == k
        rep["levels"][str(k)] = {
            "vertices": int(m.sum()),
            "relative_error": float(np.linalg.norm(e[m]) / max(np.linalg.norm(w_star[m]), 1e-300)),
            "share_of_error": float(np.linalg.norm(e[m]) ** 2 / np.linalg.norm(e) ** 2),
            "error_on_this_level_only": start(w_star + np.where(m, e, 0.0)),
            "error_removed_on_this_level": start(w_net - np.where(m, e, 0.0)),
        }
    svo.close()
    return rep


def summarize(reps):
    def mean(path):
        vals = []
        for r in reps:
            v = r
            for p in path:
                v = v[p]
            vals.append(v)
        return float(np.mean(vals))

    out = {"meshes": len(reps), "mean_n": mean(["n"]), "iterations_cold": mean(["cold", "iterations"]),
           "iterations_network": mean(["network", "iterations"]),
           "initial_normal_rel_network": mean(["network", "initial_normal_rel"]),
           "iterations_exact_dual": mean(["exact_dual", "iterations"]),
           "network_relative_error": mean(["relative_error"]),
           "scaled_network_error": {}, "scaled_iid_error": {}, "levels": {}}
    for sc in SCALES:
        k = str(sc)
        out["scaled_network_error"][k] = {"iterations": mean(["scaled_network_error", k, "iterations"]),
                                          "initial_normal_rel": mean(["scaled_network_error", k,
                                                                      "initial_normal_rel"])}
        out["scaled_iid_error"][k] = {"iterations": mean(["scaled_iid_error", k, "iterations"]),
                                      "initial_normal_rel": mean(["scaled_iid_error", k, "initial_normal_rel"])}
    for k in reps[0]["levels"]:
        if not all(k in r["levels"] for r in reps):
            continue
        out["levels"][k] = {"vertices": mean(["levels", k, "vertices"]),
                            "relative_error": mean(["levels", k, "relative_error"]),
                            "share_of_error": mean(["levels", k, "share_of_error"]),
                            "iterations_error_on_this_level_only": mean(["levels", k, "error_on_this_level_only",
                                                                         "iterations"]),
                            "iterations_error_removed_on_this_level": mean(["levels", k,
                                                                            "error_removed_on_this_level",
                                                                            "iterations"])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=8)
    ap.add_argument("--R", type=int, default=256)
    ap.add_argument("--weights", default=os.path.join(ROOT, "its_amd", "data", "warm_mlp.npz"))
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.cuda.init()
    from its_amd import _capi
    ctx = _capi.Context(0)
    net = _capi.Mlp(ctx)
    net.set_params(np.load(a.weights)["params"])
    reps = []
    for k in range(a.meshes):
        reps.append(study_mesh(ctx, torch, net, 500000 + k, a.R))
        r = reps[-1]
        log(f"[warm report] mesh {k}: n={r['n']} cold {r['cold']['iterations']} network {r['network']['iterations']} "
            f"exact {r['exact_dual']['iterations']}; scaled error "
            f"{[r['scaled_network_error'][str(s)]['iterations'] for s in SCALES]}")
    summary = summarize(reps)
    print(json.dumps(summary))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump({"summary": summary, "per_mesh": reps}, fh, indent=1)
    net.close()


if __name__ == "__main__":
    main()
