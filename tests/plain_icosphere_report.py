"""Plain method, small: the whole path on the GPU beside the CPU restatement.

Workload: icosphere subdivided 4 times (2,562 vertices, 5,120 triangles)
with seeded radial noise r = 1 + 0.05 * sum a_lm Y_lm over the 25 real
spherical harmonics l <= 4, a_lm ~ N(0, 1), normalised to the unit cube;
octree at R = 64 (7 levels); fit to a relative normal residual of 1e-6;
f at all mesh vertices and on the 64^3 marching-cubes grid.

Each stage runs twice, timed by wall clock with inputs and outputs in host
memory: through the library (octree build, signed distance at the octree
vertices, binary32/binary64 LSCG, binary32 evaluation, marching cubes), and
through the binary64 CPU restatement of the reference (SparseVoxelOctree::
createOctree, the signed-distance right-hand side, Eigen's LSCG over the
assembled matrix, ThinShells::getBSplineValue at the mesh vertices,
MC::marching_cubes) on the host's cores.  Reported: the octree's bit
equality, and max-abs / relative errors of b, lambda, f at the mesh
vertices and the marching-cubes soup against the restatement.

    python tests/plain_icosphere_report.py [--seed 1] [--R 64] [--out report.json]

Also run at this size by tests/test_gpu_plain.py.
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


def _err(a, ref):
    a, ref = np.asarray(a, np.float64), np.asarray(ref, np.float64)
    if a.shape != ref.shape:
        return {"shape": [list(a.shape), list(ref.shape)]}
    m = float(np.abs(a - ref).max()) if a.size else 0.0
    return {"max_abs": m, "rel_max": m / max(float(np.abs(ref).max()) if ref.size else 0.0, 1e-300)}


def _svo_equal(gpu, ref):
    """Every octree field, bit for bit (binary64 geometry compared as bits)."""
    if (gpu.num_nodes, gpu.num_verts, gpu.depth) != (ref.num_nodes, ref.num_verts, ref.depth):
        return False
    nd = gpu.nodes()
    xyz, owner = gpu.verts()
    pairs = [(nd["morton"], ref.morton), (nd["is_leaf"], ref.is_leaf), (nd["parent"], ref.parent),
             (nd["childs"], ref.childs), (nd["neighbors"], ref.neighbors),
             (nd["origin"].view(np.uint64), ref.node_origin.view(np.uint64)),
             (nd["width"].view(np.uint64), ref.node_width.view(np.uint64)),
             (owner, ref.vert_owner), (xyz.view(np.uint64), ref.vert_xyz.view(np.uint64))]
    return all(np.array_equal(a, b) for a, b in pairs)


def run(ctx, seed=1, R=64, tol=1e-6):
    import oracle as O
    from its_amd import _capi, meshes
    V0, F = meshes.noisy_icosphere(seed=seed, level=4)
    F = np.ascontiguousarray(F, np.int32)
    V = O.normalize(V0)
    origin, end, width = O.uniform_bbox(V)
    tris = meshes.triangles(V, F)
    res = (R, R, R)
    rep = {"workload": f"noisy icosphere level 4 (seed {seed}): {len(V)} vertices, {len(F)} triangles, R={R}, "
                       f"LSCG to {tol:g}, f at the mesh vertices and on the {R}^3 marching-cubes grid",
           "cpu_threads": int(O.lib().ora_num_threads())}
    tg, tc = {}, {}

    def gpu_timed(key, fn):
        ctx.synchronize()
        t = time.perf_counter()
        out = fn()
        ctx.synchronize()
        tg[key] = time.perf_counter() - t
        return out

    def cpu_timed(key, fn):
        t = time.perf_counter()
        out = fn()
        tc[key] = time.perf_counter() - t
        return out

    # library path (a first throwaway build loads the code objects)
    _capi.Octree(ctx, tris, origin, width, R).close()
    g = gpu_timed("octree", lambda: _capi.Octree(ctx, tris, origin, width, R))
    b = gpu_timed("rhs", lambda: g.sdf_at_vertices(V, F))
    lam, st = gpu_timed("lscg", lambda: g.lscg(b, tol=tol))
    gpu_timed("set_lambda", lambda: g.set_lambda(lam))
    f_v = gpu_timed("eval_vertices", lambda: g.eval(V, end))
    tri, mst = gpu_timed("mc", lambda: g.marching_cubes(origin, width, res, 0.0))
    log(f"[plain] gpu: n={g.num_verts}, {st['iterations']} iterations, {len(tri)} soup vertices, "
        + ", ".join(f"{k} {v * 1e3:.2f} ms" for k, v in tg.items()))

    # binary64 restatement on the host
    ref = cpu_timed("octree", lambda: O.SVO(tris, origin, width, R))
    b_r = cpu_timed("rhs", lambda: O.signed_distance(V, F, ref.vert_xyz))
    A = cpu_timed("assemble", lambda: O.Matrix(ref))
    lam_r, st_r = cpu_timed("lscg", lambda: A.lscg(b_r, tol=tol))
    f_r = cpu_timed("eval_vertices", lambda: ref.eval_support(lam_r, V, end, 0.0))
    soup_r, cube_r, nvalid_r = cpu_timed("mc", lambda: O.marching_cubes(ref, lam_r, origin, width, res, 0.0))
    log("[plain] cpu: " + ", ".join(f"{k} {v * 1e3:.1f} ms" for k, v in tc.items()))
    lam_t, _ = A.lscg(b_r, x0=lam_r, tol=1e-12)

    rep.update({"n": int(g.num_verts), "nodes": int(g.num_nodes), "depth": int(g.depth),
                "svo_bit_exact": bool(_svo_equal(g, ref)),
                "iterations": {"gpu": st["iterations"], "cpu": st_r["iterations"]},
                "normal_rel": {"gpu": st["normal_rel"], "cpu": st_r["normal_rel"]}})
    rep["b"] = _err(b, b_r)
    rep["lambda_vs_cpu_same_tol"] = _err(lam, lam_r)
    rep["lambda_vs_converged"] = _err(lam, lam_t)
    rep["cpu_lambda_vs_converged"] = _err(lam_r, lam_t)
    rep["f_vertices"] = _err(f_v, f_r)
    # evaluation alone: the library's binary32 field of the restatement's lambda
    g.set_lambda(lam_r)
    rep["f_vertices_same_lambda"] = _err(g.eval(V, end), f_r)
    tri_same, _ = g.marching_cubes(origin, width, res, 0.0)
    h = width[0] / R
    same = {"soup_vertices": [int(len(tri_same)), int(len(soup_r))]}
    if len(tri_same) == len(soup_r):
        d = float(np.abs(tri_same.astype(np.float64) - soup_r).max()) if len(soup_r) else 0.0
        same.update({"vertex_max_abs": d, "vertex_max_in_voxels": d / h})
    rep["mc_same_lambda"] = same
    rep["mc"] = {"soup_vertices": {"gpu": int(mst["num_verts"]), "cpu": int(len(soup_r))},
                 "valid_voxels": {"gpu": int(mst["num_valid_voxels"]), "cpu": int(nvalid_r)}}
    if len(tri) == len(soup_r):
        d = float(np.abs(tri.astype(np.float64) - soup_r).max()) if len(soup_r) else 0.0
        rep["mc"].update({"vertex_max_abs": d, "vertex_max_in_voxels": d / h})
    rep["seconds_gpu"] = tg
    rep["seconds_cpu"] = tc
    rep["speedup"] = {k: tc[k] / tg[k] for k in tg if k in tc and tg[k] > 0}
    rep["speedup"]["total"] = sum(tc.values()) / sum(tg.values())
    g.close()
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.cuda.init()
    from its_amd import _capi
    ctx = _capi.Context(0)
    rep = run(ctx, a.seed, a.R)
    print(json.dumps(rep))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)


if __name__ == "__main__":
    main()
