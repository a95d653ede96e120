"""Precision report: binary32 evaluation and matvec with binary64 refinement,
against a binary64 fit.

Workload (precision study): union of 32 seeded spheres (centres
U[0.2,0.8]^3, radii U[0.05,0.15]) remeshed by marching cubes to about 1.5 M
triangles, normalised to the unit cube, octree at R = 512 (10 levels).

The product path (its_lscg_solve: binary32 CGLS with binary64 dots,
residuals and refinement; binary32 evaluation) is compared with a binary64
fit: Eigen's LSCG algorithm run in binary64 (oracle.lscg_operator) over the
library's binary64 collocation products, which tests/test_lscg pins to the
assembled oracle matrix (<= 1e-12).  Reported: max-abs and relative error of
lambda, of f at the octree vertices (binary32 evaluation vs A lambda_64) and of
the marching-cubes vertices on the octree's level-0 lattice (iso 0), all in
one JSON document.

    python tests/precision_report.py [--R 512] [--tris 1.5e6] [--out report.json]

Also run at small size by tests/test_precision.py.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def blob_mesh(seed, target_tris, device=None):
    from its_amd import grid_mc
    V, F = grid_mc.blob_union(seed, res=64, device=device)
    res = max(16, int(round(64 * np.sqrt(target_tris / len(F)))))
    return grid_mc.blob_union(seed, res=res, device=device), res


def run(ctx, R=512, target_tris=1.5e6, seed=42, tol=1e-6, device=None):
    import oracle as O
    from its_amd import _capi, grid_mc, meshes
    t0 = time.time()
    (V, F), mres = blob_mesh(seed, target_tris, device)
    V = _capi.model_to_unit_cube(V)
    origin, end, width = _capi.uniform_bbox(V)
    log(f"[precision] blob mesh: {len(F)} triangles (grid {mres}^3), {time.time() - t0:.1f} s")
    svo = _capi.Octree(ctx, meshes.triangles(V, F), origin, width, R)
    n = svo.num_verts
    t1 = time.time()
    b = svo.sdf_at_vertices(V, F)
    log(f"[precision] octree R={R}: n={n}, depth {svo.depth}; signed distance {time.time() - t1:.1f} s")
    # product path
    lam_m, st_m = svo.lscg(b, tol=tol)
    log(f"[precision] mixed-precision fit: {st_m['iterations']} iterations, {st_m['seconds']:.3f} s, "
        f"normal_rel {st_m['normal_rel']:.2e}")
    # binary64 fit over the binary64 operator
    invd = svo.lscg_diag()
    t2 = time.time()
    lam_64, st_64 = O.lscg_operator(svo.mul, lambda r: svo.mul(r, transpose=True), invd, b, tol=tol)
    log(f"[precision] binary64 fit: {st_64['iterations']} iterations, {time.time() - t2:.1f} s")
    lam_64t, st_64t = O.lscg_operator(svo.mul, lambda r: svo.mul(r, transpose=True), invd, b, tol=1e-10,
                                      x0=lam_64)
    log(f"[precision] binary64 fit to 1e-10: {st_64t['iterations']} more iterations")

    def err(a, ref):
        a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
        m = float(np.abs(a - ref).max())
        return {"max_abs": m, "rel_max": m / max(float(np.abs(ref).max()), 1e-300),
                "rel_l2": float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))}

    rep = {"workload": f"blob union (32 spheres, seed {seed}) remeshed to {len(F)} triangles, R={R}",
           "n": n, "triangles": int(len(F)), "depth": svo.depth, "tol": tol,
           "iterations_mixed": st_m["iterations"], "iterations_f64": st_64["iterations"],
           "normal_rel_mixed": st_m["normal_rel"], "normal_rel_f64": st_64["normal_rel"]}
    rep["lambda_vs_f64_same_tol"] = err(lam_m, lam_64)
    rep["lambda_vs_f64_converged"] = err(lam_m, lam_64t)
    # f at the vertices: binary32 evaluation of the product field vs A lambda_64
    svo.set_lambda(lam_m)
    xyz, _ = svo.verts()
    f32 = svo.eval(xyz, end, _capi.EVAL_RAW)
    f64 = svo.mul(lam_64)
    f64t = svo.mul(lam_64t)
    rep["f_vertices_vs_f64_same_tol"] = err(f32, f64)
    rep["f_vertices_vs_f64_converged"] = err(f32, f64t)
    rep["f_vertices_vs_b"] = err(f32, b)
    # marching cubes on the level-0 lattice, iso 0
    lat32 = svo.mc_lattice(origin, width, (R, R, R)).astype(np.float64).reshape(R + 1, R + 1, R + 1)
    tri, mst = svo.marching_cubes(origin, width, (R, R, R), 0.0)
    svo.set_lambda(lam_64t)
    lat64 = svo.lattice_f64().reshape(R + 1, R + 1, R + 1)
    h = width / R
    soup32, case32 = grid_mc.marching_cubes(lat32, origin, h, 0.0, weld=False, device=device)
    soup64, case64 = grid_mc.marching_cubes(lat64, origin, h, 0.0, weld=False, device=device)
    # the library's own soup is the binary32-lattice soup rounded to binary32
    same_lib = len(tri) == len(soup32) and float(np.abs(tri - soup32).max()) <= 1e-6
    nvt = (grid_mc.tri_table() >= 0).sum(1)
    both = case32 == case64
    # vertices of the cubes whose case agrees, matched in soup order
    cnt32, cnt64 = nvt[case32], nvt[case64]
    off32 = np.concatenate([[0], np.cumsum(cnt32)])
    off64 = np.concatenate([[0], np.cumsum(cnt64)])
    cubes = np.nonzero(both & (cnt32 > 0))[0]
    k = cnt32[cubes]
    within = np.arange(int(k.sum())) - np.repeat(np.cumsum(k) - k, k)
    idx32 = np.repeat(off32[cubes], k) + within
    idx64 = np.repeat(off64[cubes], k) + within
    d = np.abs(soup32[idx32] - soup64[idx64]).max(1) if len(idx32) else np.zeros(1)
    rep["mc"] = {"res": R, "vertices_mixed": int(len(soup32)), "vertices_f64": int(len(soup64)),
                 "library_soup_equals_recomputed": bool(same_lib),
                 "cubes_with_other_case": int((~both).sum()), "matched_vertices": int(len(idx32)),
                 "vertex_max_abs": float(d.max()), "vertex_max_in_voxels": float(d.max() / h[0]),
                 "vertex_rms": float(np.sqrt(np.mean(d * d)))}
    rep["seconds"] = time.time() - t0
    svo.close()
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--R", type=int, default=512)
    ap.add_argument("--tris", type=float, default=1.5e6)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.cuda.init()
    from its_amd import _capi
    ctx = _capi.Context(0)
    rep = run(ctx, a.R, a.tris, a.seed)
    print(json.dumps(rep))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)


if __name__ == "__main__":
    main()
