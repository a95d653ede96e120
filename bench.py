"""Benchmark: B-spline evals/s at depth 8 (plus LSCG time to 1e-6).

Workload: the (3,5) torus-knot tube with 200,000 triangles and seeded vertex
jitter N(0, 1e-3), normalised to the unit cube.  The octree has depth 8
(256^3 voxels, 9 levels).  The coefficients come from the GPU LSCG fit of
the signed distance at the node vertices.  There are 10^7 binary32 query
points, uniform in [-0.05, 1.05]^3, already resident in HBM.

A step is one evaluation of f at all 10^7 points (its_bspline_eval_device,
support semantics of the reference CPU query path).  The timed steps take
four query/output buffer sets in turn (640 MB, more than the 256 MB Infinity
Cache), so a launch does not find the previous launch's data in that cache;
the same launches on one buffer set are timed and reported beside them.  The JSON line also
carries:
- the cold-start LSCG time to a 1e-6 normal residual on the same octree;
- marching cubes on the 256^3 grid of the octree box (iso 0);
- the learned warm start against the cold start on held-out noisy
  icospheres at R = 256 (weights its_amd/data/warm_mlp.npz);
- the roofline of the evaluation kernel (algorithmic bytes / measured launch
  time);
- the oracle CPU path, timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_QUERIES = 10_000_000
K_SETS = 4
R = 256
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_mesh():
    from its_amd import _capi, meshes

    V, F = meshes.torus_knot_tube()  # 2000 x 50 x 2 = 200,000 triangles
    V = _capi.model_to_unit_cube(V)
    V = meshes.jitter(V, 1e-3, seed=7)
    origin, end, width = _capi.uniform_bbox(V)
    return V, F, origin, end, width


def cpu_baseline(V, F, origin, end, width, lam, q_host, f_gpu, gpu_svo, budget_s=15.0):
    """Oracle (FP64 CPU restatement of mt_cpuTest_2) on a prefix of the
    queries, timed; its values double as the parity check of the timed GPU
    output, and the oracle's octree is compared with the GPU's bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from its_amd import meshes

    svo = O.SVO(meshes.triangles(V, F), origin, width, R)
    nd = gpu_svo.nodes()
    xyz, owner = gpu_svo.verts()
    bit_exact = bool(svo.num_nodes == gpu_svo.num_nodes and svo.num_verts == gpu_svo.num_verts
                     and all(np.array_equal(nd[k], r) for k, r in
                             [("morton", svo.morton), ("is_leaf", svo.is_leaf), ("parent", svo.parent),
                              ("childs", svo.childs), ("neighbors", svo.neighbors)])
                     and np.array_equal(nd["origin"].view(np.uint64), svo.node_origin.view(np.uint64))
                     and np.array_equal(nd["width"].view(np.uint64), svo.node_width.view(np.uint64))
                     and np.array_equal(owner, svo.vert_owner)
                     and np.array_equal(xyz.view(np.uint64), svo.vert_xyz.view(np.uint64)))
    threads = O.lib().ora_num_threads()
    n = 20000
    t0 = time.perf_counter()
    svo.eval_support(lam, q_host[:n].astype(np.float64), end, 0.0)
    dt = time.perf_counter() - t0
    n2 = int(min(len(q_host), max(n, budget_s / max(dt, 1e-9) * n)))
    t0 = time.perf_counter()
    f_ref = svo.eval_support(lam, q_host[:n2].astype(np.float64), end, 0.0)
    dt2 = time.perf_counter() - t0
    d = np.abs(f_gpu[:n2].astype(np.float64) - f_ref)
    parity = {"svo_bit_exact": bit_exact, "nodes": int(svo.num_nodes), "verts": int(svo.num_verts),
              "f_queries_checked": n2, "f_max_abs": float(d.max()),
              "f_max_rel": float((d / np.maximum(np.abs(f_ref), 1e-3)).max()), "f_tol": 1e-5,
              "f_within_tol": bool(d.max() <= 1e-5),
              "note": "timed GPU output vs the FP64 oracle support evaluation with the same coefficients; "
                      "relative error against max(|f|, 1e-3)"}
    cpu = {
        "value": n2 / dt2,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n2} of the {N_QUERIES} queries, oracle support evaluation (FP64, OpenMP)",
    }
    return cpu, parity


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--warm-meshes", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB device to device (a dispatch of known size "
                         "for calibrating hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one MI355X per box).", file=sys.stderr)
        sys.exit(2)

    import torch

    from its_amd import _capi

    dev = torch.device("cuda:0")
    torch.cuda.init()
    # one stream (default priority) for torch and the library, so the events
    # bracket exactly the kernels of the step
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = _capi.Context(0)
    ctx.set_stream(stream.cuda_stream)
    # launch shapes recorded by the autotuner for this depth
    from its_amd import autotune
    tuned_depth = autotune.apply(R.bit_length() - 1)
    tuning = {"table": "its_amd/data/autotune.json", "depth": tuned_depth,
              **{k: _capi.tune_get(v) for k, v in (("dense_cells", _capi.TUNE_DENSE_CELLS),
                                                   ("eval_blocks_per_cu", _capi.TUNE_EVAL_BLOCKS_PER_CU),
                                                   ("eval_ilp", _capi.TUNE_EVAL_ILP),
                                                   ("eval_queue", _capi.TUNE_EVAL_QUEUE),
                                                   ("coarse_verts", _capi.TUNE_COARSE_VERTS))}}
    log(f"[bench] tuning {tuning}")

    log("[bench] mesh")
    V, F, origin, end, width = make_mesh()
    from its_amd import meshes

    tris = meshes.triangles(V, F)
    t0 = time.perf_counter()
    svo = _capi.Octree(ctx, tris, origin, width, R)
    t_svo = time.perf_counter() - t0
    log(f"[bench] octree depth={svo.depth} nodes={svo.num_nodes} verts={svo.num_verts} ({t_svo:.2f}s)")
    t0 = time.perf_counter()
    b = svo.sdf_at_vertices(V, F)
    t_sdf = time.perf_counter() - t0
    log(f"[bench] rhs ({t_sdf:.2f}s)")
    # the same two steps again, warm (code objects loaded, allocator grown):
    # median of 3, host clock, triangles and points in host memory
    warm_svo, warm_sdf = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        s2 = _capi.Octree(ctx, tris, origin, width, R)
        warm_svo.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        s2.sdf_at_vertices(V, F)
        warm_sdf.append(time.perf_counter() - t0)
        s2.close()
    t_svo_warm, t_sdf_warm = float(np.median(warm_svo)), float(np.median(warm_sdf))
    log(f"[bench] warm: octree {t_svo_warm * 1e3:.2f} ms, rhs {t_sdf_warm * 1e3:.2f} ms")
    # warm-up fit builds the operator structure and the solver graph
    svo.lscg(b, tol=1e-6)
    lscg_runs = []
    for _ in range(3):
        lam, st = svo.lscg(b, tol=1e-6)
        lscg_runs.append(st)
    t_lscg = float(np.median([s["seconds"] for s in lscg_runs]))
    st = lscg_runs[-1]
    log(f"[bench] lscg cold: {t_lscg*1e3:.2f} ms, iters={st['iterations']} normal_rel={st['normal_rel']:.2e}")
    # LSCG roofline: bytes of one CGLS iteration over the time of one
    # iteration (the whole solve, set-up and binary64 refinement included,
    # divided by the iteration count): the matrix-free operator's own count,
    # and the assembled matrix's 16 B per non-zero + 72 B per row
    ls = svo.lscg_structure()
    t_it = t_lscg / max(st["iterations"], 1)
    lscg_roof = {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "iterations": st["iterations"],
                 "seconds_per_iteration": t_it, "nnz": ls["nnz"], "n": ls["n"],
                 "parent_entries": ls["parent_entries"],
                 "bytes_matrix_free_per_iteration": ls["bytes_matrix_free_per_iteration"],
                 "achieved": ls["bytes_matrix_free_per_iteration"] / t_it / 1e9,
                 "frac": ls["bytes_matrix_free_per_iteration"] / t_it / 1e9 / HBM_PEAK_GBS,
                 "bytes_assembled_per_iteration": ls["bytes_assembled_per_iteration"],
                 "assembled_equivalent": ls["bytes_assembled_per_iteration"] / t_it / 1e9,
                 "frac_assembled_equivalent": ls["bytes_assembled_per_iteration"] / t_it / 1e9 / HBM_PEAK_GBS}
    # measured HBM traffic of one iteration (rocprofv3 counters, summed over
    # the iteration's kernels)
    lpmc = os.path.join(ROOT, "profiles", "lscg_knot256_traffic.json")
    if os.path.exists(lpmc):
        with open(lpmc) as fh:
            lscg_roof["traffic"] = json.load(fh).get("bytes_per_iteration")
    log(f"[bench] lscg roofline: {t_it * 1e6:.1f} us per iteration, matrix-free "
        f"{lscg_roof['achieved']:.0f} GB/s, assembled-equivalent {lscg_roof['assembled_equivalent']:.0f} GB/s")

    # marching cubes on the 256^3 grid of the octree box
    # the coefficients are set again before each run, so each run evaluates
    # its lattice (the octree keeps the last one for the same coefficients)
    mc_runs = []
    for _ in range(3):
        svo.set_lambda(lam)
        mc_runs.append(svo.marching_cubes(origin, width, (R, R, R), 0.0)[1])
    mc = {"res": R, "vertices": mc_runs[-1]["num_verts"],
          "seconds_lattice": float(np.median([m["seconds_eval"] for m in mc_runs])),
          "seconds_extract": float(np.median([m["seconds_extract"] for m in mc_runs])),
          "seconds_lattice_device": float(np.median([m["seconds_eval_device"] for m in mc_runs])),
          "seconds_extract_device": float(np.median([m["seconds_extract_device"] for m in mc_runs]))}
    # algorithmic bytes: the (R+1)^3 binary32 lattice written and read, 4 B
    # per voxel of case/count, 36 B per triangle of binary32 soup
    mc_bytes = 8.0 * (R + 1) ** 3 + 4.0 * R ** 3 + 12.0 * mc["vertices"]
    mc_dev = mc["seconds_lattice_device"] + mc["seconds_extract_device"]
    mc["roofline"] = {"bound": "hbm", "algorithmic_bytes": mc_bytes, "achieved": mc_bytes / mc_dev / 1e9,
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": mc_bytes / mc_dev / 1e9 / HBM_PEAK_GBS}
    log(f"[bench] mc 256^3: {mc['vertices']} vertices, lattice {mc['seconds_lattice']*1e3:.2f} ms "
        f"(device {mc['seconds_lattice_device']*1e3:.3f}), extract {mc['seconds_extract']*1e3:.2f} ms "
        f"(device {mc['seconds_extract_device']*1e3:.3f})")

    # queries resident in HBM: K_SETS sets of queries and outputs (640 MB in
    # all), used in turn, so that a launch cannot find its queries, outputs or
    # most of the coefficient blocks left in the 256 MB Infinity Cache by the
    # launch before; set 0 is also timed alone (every launch on the same
    # buffers) for comparison
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    qs = [torch.rand((3, N_QUERIES), generator=g, device=dev, dtype=torch.float32) * 1.1 - 0.05
          for _ in range(K_SETS)]
    outs = [torch.empty(N_QUERIES, device=dev, dtype=torch.float32) for _ in range(K_SETS)]
    q, out = qs[0], outs[0]

    def step_on(k):
        qk = qs[k]
        svo.eval_device(qk[0].data_ptr(), qk[1].data_ptr(), qk[2].data_ptr(), N_QUERIES, end, outs[k].data_ptr(),
                        _capi.EVAL_SUPPORT, 0.0, stream=stream.cuda_stream)

    def timed(sets):
        for w in range(args.warmup):
            step_on(sets[w % len(sets)])
        torch.cuda.synchronize()
        # one event pair on the launch stream brackets the K back-to-back
        # launches: the average launch duration, with no event work between
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for k in range(args.steps):
            step_on(sets[k % len(sets)])
        e1.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, e0.elapsed_time(e1) / args.steps

    elapsed, avg_ms = timed(list(range(K_SETS)))
    value = N_QUERIES * args.steps / elapsed
    elapsed_1, avg_ms_1 = timed([0])
    log(f"[bench] eval: {avg_ms * 1e3:.1f} us per launch over {K_SETS} buffer sets, "
        f"{avg_ms_1 * 1e3:.1f} us on one set")
    # every set's output against a fresh evaluation of its queries
    sets_agree = True
    for k in range(1, K_SETS):
        chk = torch.empty_like(outs[k])
        svo.eval_device(qs[k][0].data_ptr(), qs[k][1].data_ptr(), qs[k][2].data_ptr(), N_QUERIES, end,
                        chk.data_ptr(), _capi.EVAL_SUPPORT, 0.0, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        sets_agree &= bool(torch.equal(chk, outs[k]))

    # end to end with the queries in host memory (outside the timed region):
    # pinned SoA buffers copied in and out on the launch stream around the
    # same launch, and the host entry point (pageable binary32 xyz in
    # interleaved order, f out)
    qh = q.cpu().pin_memory()
    fh = torch.empty(N_QUERIES, dtype=torch.float32).pin_memory()
    qd = torch.empty_like(q)

    def step_pinned():
        qd.copy_(qh, non_blocking=True)
        svo.eval_device(qd[0].data_ptr(), qd[1].data_ptr(), qd[2].data_ptr(), N_QUERIES, end, out.data_ptr(),
                        _capi.EVAL_SUPPORT, 0.0, stream=stream.cuda_stream)
        fh.copy_(out, non_blocking=True)

    def wall(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    t_pinned = wall(step_pinned)
    pinned_ok = bool(torch.equal(fh, out.cpu()))
    q_aos = qh.numpy().T.copy()
    t_host = wall(lambda: svo.eval(q_aos, end))
    host_ok = bool(np.array_equal(svo.eval(q_aos, end), out.cpu().numpy()))
    e2e = {"pinned_soa_ms": t_pinned * 1e3, "pinned_evals_per_s": N_QUERIES / t_pinned,
           "host_entry_ms": t_host * 1e3, "host_entry_evals_per_s": N_QUERIES / t_host,
           "pinned_equals_resident": pinned_ok, "host_entry_equals_resident": host_ok}
    log(f"[bench] end to end: pinned {t_pinned * 1e3:.2f} ms, host entry point {t_host * 1e3:.2f} ms")
    del qh, fh, qd, q_aos

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib * (1 << 18), device=dev, dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # achievable bandwidth on this device: the library's own streaming copy
    # of 1 GiB (read + write), median of 10 launches, after the timed region
    stream_gbs = ctx.hbm_bandwidth(1 << 30, 10)
    log(f"[bench] streaming copy: {stream_gbs:.0f} GB/s (read + write)")
    mc["roofline"]["measured_peak"] = stream_gbs
    mc["roofline"]["frac_of_measured"] = mc["roofline"]["achieved"] / stream_gbs

    # roofline of the evaluation kernel: 16 B/query (12 B of binary32 xyz in,
    # 4 B of f out) + 8 B per unknown (one pass over the coefficient field)
    alg_bytes = 16.0 * N_QUERIES + 8.0 * svo.num_verts
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "eval_knot256_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    warm = None
    wfile = os.path.join(ROOT, "its_amd", "data", "warm_mlp.npz")
    if os.path.exists(wfile) and args.warm_meshes > 0:
        from its_amd import warm_train
        net = _capi.Mlp(ctx)
        net.set_params(np.load(wfile)["params"])
        qw = np.random.default_rng(12345).uniform(0, 1, size=(20000, 3))
        rs = [warm_train.evaluate(ctx, torch, net, 500000 + k, 256, 1e-6, qw) for k in range(args.warm_meshes)]
        net.close()
        tc, tw = sum(r["t_cold"] for r in rs), sum(r["t_warm"] for r in rs)
        warm = {"workload": "held-out noisy icospheres (level 4, 5,120 triangles), R=256, tol 1e-6",
                "meshes": len(rs), "mean_iters_cold": float(np.mean([r["iters_cold"] for r in rs])),
                "mean_iters_warm": float(np.mean([r["iters_warm"] for r in rs])),
                "mean_ms_cold": tc / len(rs) * 1e3, "mean_ms_warm": tw / len(rs) * 1e3,
                "mean_ms_predict": float(np.mean([r["t_predict"] for r in rs])) * 1e3,
                "speedup": tc / tw, "target_speedup": 3.0,
                "max_f_diff_warm_vs_cold": float(max(r["f_maxdiff"] for r in rs))}
        log(f"[bench] warm start: {warm['mean_iters_cold']:.1f} -> {warm['mean_iters_warm']:.1f} iterations, "
            f"speedup {warm['speedup']:.2f}")

    cpu = parity = None
    if not args.no_cpu:
        log("[bench] cpu baseline and parity")
        cpu, parity = cpu_baseline(V, F, origin, end, width, lam, q.cpu().numpy().T.copy(), out.cpu().numpy(), svo)
        log(f"[bench] parity: octree bit-exact {parity['svo_bit_exact']}, f max abs {parity['f_max_abs']:.2e} "
            f"over {parity['f_queries_checked']} queries")

    res = {
        "metric": "B-spline evals/s @depth 8; LSCG time-to-1e-6 residual, warm vs cold start",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "torus-knot (3,5) tube, 200k triangles, jitter N(0,1e-3), SVO depth 8 (256^3), "
                        "10^7 FP32 queries uniform in [-0.05,1.05]^3",
            "queries": N_QUERIES,
            "svo_R": R,
            "tree_depth": svo.depth,
            "unknowns": svo.num_verts,
            "nodes": svo.num_nodes,
        },
        "tuning": tuning,
        "eval_kernel_ms": {"mean": avg_ms, "buffer_sets": K_SETS, "mean_one_set": avg_ms_1,
                           "note": "mean is over launches that take the query/output sets in turn (640 MB, more "
                                   "than the 256 MB Infinity Cache); mean_one_set repeats set 0",
                           "sets_reproduce": sets_agree},
        "eval_end_to_end": e2e,
        "lscg_cold": {"seconds": t_lscg, "iterations": st["iterations"], "normal_rel": st["normal_rel"],
                      "residual_rel": st["residual_rel"], "roofline": lscg_roof},
        "lscg_warm": warm,
        "mc": mc,
        "setup_seconds": {"svo_build": t_svo_warm, "sdf_rhs": t_sdf_warm, "svo_build_first_call": t_svo,
                          "sdf_rhs_first_call": t_sdf},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "measured_peak": stream_gbs,
                     "frac_of_measured": achieved / stream_gbs, "algorithmic_bytes": alg_bytes,
                     "frac_one_set": alg_bytes / (avg_ms_1 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "parity": parity,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
