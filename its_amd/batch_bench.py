"""Batching workload: many meshes of uneven size built, fitted and evaluated.

Workload (batching): 256 seeded meshes -- noisy icospheres, tori and
blob unions in turn -- with triangle counts log-uniform in [1k, 200k] and
octree depth 5-8 (R = 32..256) drawn per mesh; 65,536 uniform queries per mesh.

Stages: octree build, signed-distance right-hand side and the solver's
lists and preconditioner per mesh, from several host threads with a
context each (--build-threads); the fit of
all meshes (its_batch_fit: the smaller meshes in one persistent
graph-captured launch, the larger ones beside it through the many-mesh
solver, one launch per solver pass for all of them); the evaluation fields
with the autotuned collapse bound of each depth (its_batch_fields);
the evaluation of all meshes' queries in one launch (its_batch_eval_device).
The same fits and evaluations also run mesh by mesh through the single-mesh
entry points for comparison.  Prints one JSON line.

    python -m its_amd.batch_bench [--meshes 256] [--queries 65536] [--build-threads 12] [--caps 1,65536]
"""
import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_mesh(k, rng):
    from its_amd import grid_mc, meshes
    T = math.exp(rng.uniform(math.log(1e3), math.log(2e5)))
    kind = ("icosphere", "torus", "blob")[k % 3]
    if kind == "icosphere":
        level = int(np.clip(round(math.log(T / 20.0, 4)), 3, 6))
        V, F = meshes.noisy_icosphere(seed=k, level=level)
    elif kind == "torus":
        nv = max(8, int(math.sqrt(T / 6.0)))
        nu = max(8, int(T / (2 * nv)))
        V, F = meshes.torus(nu, nv)
    else:
        V0, F0 = grid_mc.blob_union(k, res=32)
        res = int(np.clip(round(32 * math.sqrt(T / max(len(F0), 1))), 12, 400))
        V, F = grid_mc.blob_union(k, res=res)
    depth = int(rng.integers(5, 9))
    return kind, V, F, depth


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=256)
    ap.add_argument("--queries", type=int, default=65536)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--table", default=os.path.join(HERE, "data", "autotune.json"))
    ap.add_argument("--build-threads", type=int, default=min(12, os.cpu_count() or 4),
                    help="host threads building octrees and right-hand sides, one context each (on one "
                         "MI355X the 256-mesh end-to-end rate rose from 4 to 12 threads and not beyond)")
    ap.add_argument("--caps", default="",
                    help="comma-separated ITS_TUNE_BATCH_MAX_VERTS values: the whole fit is timed once more "
                         "at each (meshes above the cap go to the many-mesh solver)")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()
    from its_amd import _capi, autotune, meshes
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = _capi.Context(0)
    ctx.set_stream(stream.cuda_stream)
    table = autotune.load(a.table) if os.path.exists(a.table) else None
    rng = np.random.default_rng(a.seed)

    t0 = time.perf_counter()
    specs = [make_mesh(k, rng) for k in range(a.meshes)]
    t_gen = time.perf_counter() - t0
    log(f"[batch] {a.meshes} meshes generated ({t_gen:.1f} s)")
    # octrees and right-hand sides, from build_threads host threads with a
    # context (stream) each, meshes taken in order from a shared counter;
    # the batch below joins octrees of any context on the device
    nthr = max(1, a.build_threads)
    bctx = [ctx] if nthr == 1 else [_capi.Context(0) for _ in range(nthr)]
    items = [None] * a.meshes
    times = np.zeros((a.meshes, 3))
    lock = threading.Lock()
    nxt = [0]
    errors = []

    def build_worker(t):
        try:
            while True:
                with lock:
                    k = nxt[0]
                    nxt[0] += 1
                if k >= a.meshes:
                    return
                kind, V, F, depth = specs[k]
                V = _capi.model_to_unit_cube(V)
                origin, end, width = _capi.uniform_bbox(V)
                s0 = time.perf_counter()
                g = _capi.Octree(bctx[t], meshes.triangles(V, F), origin, width, 1 << depth)
                s1 = time.perf_counter()
                b = g.sdf_at_vertices(V, F)
                s2 = time.perf_counter()
                g.lscg_structure()                   # the solver's lists and preconditioner
                s3 = time.perf_counter()
                times[k] = (s1 - s0, s2 - s1, s3 - s2)
                items[k] = dict(kind=kind, T=len(F), depth=depth, g=g, b=b, origin=origin, width=width)
        except Exception as e:   # re-raised below
            errors.append(e)

    w0 = time.perf_counter()
    workers = [threading.Thread(target=build_worker, args=(t,)) for t in range(nthr)]
    for th in workers:
        th.start()
    for th in workers:
        th.join()
    t_build_wall = time.perf_counter() - w0
    if errors:
        raise errors[0]
    t_build, t_rhs, t_setup = (float(times[:, c].sum()) for c in range(3))
    log(f"[batch] built {a.meshes} on {nthr} threads in {t_build_wall:.2f} s (summed per-call: octrees "
        f"{t_build:.2f} s, right-hand sides {t_rhs:.2f} s, solver setup {t_setup:.2f} s)")

    # fit: mesh by mesh, then all in one launch
    for it in items:
        it["g"].lscg(it["b"], tol=1e-6)              # solver structures and graphs
    seq_fit = 0.0
    seq_stats = []
    for it in items:
        lam, st = it["g"].lscg(it["b"], tol=1e-6)
        seq_fit += st["seconds"]
        seq_stats.append(st)
    batch = _capi.Batch(ctx)
    for it in items:
        batch.add(it["g"])
    batch.fit(tol=1e-6)                              # captures the graph
    stats, t_fit = batch.fit(tol=1e-6)
    log(f"[batch] fit: one launch {t_fit * 1e3:.1f} ms vs mesh by mesh {seq_fit * 1e3:.1f} ms")

    # evaluation fields with the collapse bound of each depth: one
    # its_batch_fields call per bound (octrees of different contexts are
    # built side by side)
    s0 = time.perf_counter()
    groups = {}
    for it in items:
        bound = (table[it["depth"]]["dense_cells"] if table and it["depth"] in table
                 else _capi.tune_get(_capi.TUNE_DENSE_CELLS))
        groups.setdefault(int(bound), []).append(it)
    old_bound = _capi.tune_get(_capi.TUNE_DENSE_CELLS)
    for bound, group in sorted(groups.items()):
        _capi.tune_set(_capi.TUNE_DENSE_CELLS, bound)
        fb = _capi.Batch(ctx)
        for it in group:
            fb.add(it["g"])
        fb.fields()
        fb.close()
    _capi.tune_set(_capi.TUNE_DENSE_CELLS, old_bound)
    t_fields = time.perf_counter() - s0
    if table:
        autotune.apply(8, table)

    # evaluation: all queries in one launch vs one launch per mesh
    nq = a.queries
    offs = np.arange(a.meshes + 1, dtype=np.int64) * nq
    g = torch.Generator(device="cuda")
    g.manual_seed(99)
    U = torch.rand((3, a.meshes * nq), generator=g, device="cuda", dtype=torch.float32)
    for m, it in enumerate(items):
        lo = torch.tensor(it["origin"] - 0.05 * it["width"], device="cuda", dtype=torch.float32)[:, None]
        wd = torch.tensor(1.1 * it["width"], device="cuda", dtype=torch.float32)[:, None]
        U[:, offs[m]:offs[m + 1]] = lo + wd * U[:, offs[m]:offs[m + 1]]
    out_b = torch.empty(a.meshes * nq, device="cuda", dtype=torch.float32)
    out_s = torch.empty_like(out_b)
    batch.eval_device(U[0].data_ptr(), U[1].data_ptr(), U[2].data_ptr(), offs, out_b.data_ptr())
    t_eval = float(np.median([batch.eval_device(U[0].data_ptr(), U[1].data_ptr(), U[2].data_ptr(), offs,
                                                out_b.data_ptr()) for _ in range(5)]))

    def seq_eval():
        for m, it in enumerate(items):
            end = it["origin"] + it["width"]
            it["g"].eval_device(U[0, offs[m]:].data_ptr(), U[1, offs[m]:].data_ptr(), U[2, offs[m]:].data_ptr(), nq,
                                end, out_s[offs[m]:].data_ptr(), _capi.EVAL_SUPPORT, 0.0, stream=stream.cuda_stream)
    seq_eval()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    seq_eval()
    e1.record(stream)
    torch.cuda.synchronize()
    t_seq_eval = e0.elapsed_time(e1) * 1e-3
    same = bool(torch.equal(out_b, out_s))
    log(f"[batch] eval: one launch {t_eval * 1e3:.2f} ms vs mesh by mesh {t_seq_eval * 1e3:.2f} ms (same {same})")

    # the fit again at other one-launch caps (best of three after a capture)
    cap_sweep = {}
    if a.caps:
        cap0 = _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS)
        for cap in [int(c) for c in a.caps.split(",") if c]:
            _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, cap)
            cb = _capi.Batch(ctx)
            for it in items:
                cb.add(it["g"])
            cb.fit(tol=1e-6)
            best = min(cb.fit(tol=1e-6)[1] for _ in range(3))
            cb.close()
            cap_sweep[cap] = {"fit_s": best, "one_launch_meshes": int(sum(1 for it in items
                                                                            if it["g"].num_verts <= cap))}
            log(f"[batch] cap {cap}: fit {best * 1e3:.1f} ms ({cap_sweep[cap]['one_launch_meshes']} meshes "
                f"in the one-launch kernel)")
        _capi.tune_set(_capi.TUNE_BATCH_MAX_VERTS, cap0)

    n_tot = sum(it["g"].num_verts for it in items)
    gpu_total = t_fit + t_eval
    res = {
        "workload": f"{a.meshes} meshes (icosphere/torus/blob union), triangles log-uniform in [1k, 200k], "
                    f"depth 5-8, {nq} queries per mesh",
        "meshes": a.meshes, "triangles_total": int(sum(it["T"] for it in items)), "unknowns_total": int(n_tot),
        "depth_counts": {d: int(sum(1 for it in items if it["depth"] == d)) for d in range(5, 9)},
        "fit_one_launch_s": t_fit, "fit_mesh_by_mesh_s": seq_fit, "fit_speedup": seq_fit / t_fit,
        "eval_one_launch_s": t_eval, "eval_mesh_by_mesh_s": t_seq_eval, "eval_speedup": t_seq_eval / t_eval,
        "eval_bit_identical": same,
        "all_converged": bool(all(s["converged"] for s in stats)),
        "max_iterations": int(max(s["iterations"] for s in stats)),
        "build_threads": nthr, "build_rhs_wall_s": t_build_wall,
        "build_s_summed": t_build, "rhs_s_summed": t_rhs, "solver_setup_s_summed": t_setup, "fields_s": t_fields,
        "meshes_per_s_fit_eval": a.meshes / gpu_total,
        "meshes_per_s_end_to_end": a.meshes / (t_build_wall + t_fit + t_fields + t_eval),
        "latency_ms_per_mesh_end_to_end": (t_build_wall + t_fit + t_fields + t_eval) / a.meshes * 1e3,
        "autotune_table": os.path.basename(a.table) if table else None,
        "one_launch_cap": _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS),
        "one_launch_meshes": int(sum(1 for it in items
                                     if it["g"].num_verts <= _capi.tune_get(_capi.TUNE_BATCH_MAX_VERTS))),
        "cap_sweep": cap_sweep,
    }
    print(json.dumps(res))
    batch.close()
    for it in items:
        it["g"].close()
    if nthr > 1:
        for c in bctx:
            c.close()
    return res


if __name__ == "__main__":
    main()
