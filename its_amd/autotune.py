"""Autotuner for the library's launch shapes, per octree depth.

For each depth (R = 2^depth voxels per axis) it builds a representative
fitted octree (the (3,5) torus-knot tube of 200,000 triangles with vertex
jitter N(0, 1e-3), seed 7), then times:
  - the evaluation kernel over 10^7 uniform queries in [-0.05, 1.05]^3,
    taking four query/output sets in turn (640 MB, more than the Infinity
    Cache), for every combination of ITS_TUNE_DENSE_CELLS (collapse bound of
    the evaluation field), ITS_TUNE_EVAL_BLOCKS_PER_CU, ITS_TUNE_EVAL_ILP and
    ITS_TUNE_EVAL_QUEUE;
  - the cold LSCG solve for each ITS_TUNE_COARSE_VERTS.
Every candidate is timed 5 times; the median decides, and ties go to the
earlier candidate in the fixed candidate order (choose()), so the choice for
a given set of timings is deterministic.  The table is written as JSON and read back by
load()/apply(), which set the process-wide tunables for a depth (the nearest
recorded depth for one that was not tuned).

    python -m its_amd.autotune --depths 5 6 7 8 --out its_amd/data/autotune.json
"""
import argparse
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
DEFAULT_TABLE = os.path.join(HERE, "data", "autotune.json")

DENSE_CELLS = (512, 4096, 32768, 262144)
BLOCKS_PER_CU = (16, 32, 64)
ILP = (1, 2)
QUEUE = (0, 1)
COARSE_VERTS = (1024, 2048, 4096, 8192)


def choose(candidates):
    """The candidate with the smallest time (its first entry); ties go to the
    earlier candidate in the list, so equal timings give the same choice."""
    best = None
    for c in candidates:
        if best is None or c[0] < best[0]:
            best = c
    return best


def _median_ms(fn, torch, stream, reps=5):
    ts = []
    for k in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn(k)
        e1.record(stream)
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1))
    return float(np.median(ts))


def tune_depth(ctx, torch, stream, depth, n_queries=10_000_000, sets=4):
    from its_amd import _capi, meshes
    R = 1 << depth
    V, F = meshes.torus_knot_tube()
    V = _capi.model_to_unit_cube(V)
    V = meshes.jitter(V, 1e-3, seed=7)
    origin, end, width = _capi.uniform_bbox(V)
    tris = meshes.triangles(V, F)
    saved = {k: _capi.tune_get(k) for k in (_capi.TUNE_DENSE_CELLS, _capi.TUNE_EVAL_BLOCKS_PER_CU,
                                            _capi.TUNE_EVAL_ILP, _capi.TUNE_EVAL_QUEUE, _capi.TUNE_COARSE_VERTS)}
    # LSCG coarse block (the solver structure is built per octree)
    lscg = []
    b = None
    for cv in COARSE_VERTS:
        _capi.tune_set(_capi.TUNE_COARSE_VERTS, cv)
        svo = _capi.Octree(ctx, tris, origin, width, R)
        if b is None:
            b = svo.sdf_at_vertices(V, F)
        else:
            svo.sdf_at_vertices(V, F)
        svo.lscg(b, tol=1e-6)
        ms = float(np.median([svo.lscg(b, tol=1e-6)[1]["seconds"] for _ in range(5)])) * 1e3
        lscg.append((ms, cv))
        lam = svo.get_lambda()
        svo.close()
    best_lscg = choose(lscg)
    _capi.tune_set(_capi.TUNE_COARSE_VERTS, best_lscg[1])
    # evaluation
    svo = _capi.Octree(ctx, tris, origin, width, R)
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    qs = [torch.rand((3, n_queries), generator=g, device="cuda", dtype=torch.float32) * 1.1 - 0.05
          for _ in range(sets)]
    outs = [torch.empty(n_queries, device="cuda", dtype=torch.float32) for _ in range(sets)]
    evals = []
    for dc in DENSE_CELLS:
        _capi.tune_set(_capi.TUNE_DENSE_CELLS, dc)
        svo.set_lambda(lam)
        for bpc, ilp, queue in itertools.product(BLOCKS_PER_CU, ILP, QUEUE):
            if queue and ilp != 1:
                continue
            _capi.tune_set(_capi.TUNE_EVAL_BLOCKS_PER_CU, bpc)
            _capi.tune_set(_capi.TUNE_EVAL_ILP, ilp)
            _capi.tune_set(_capi.TUNE_EVAL_QUEUE, queue)

            def step(k):
                q, out = qs[k % sets], outs[k % sets]
                svo.eval_device(q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(), n_queries, end, out.data_ptr(),
                                _capi.EVAL_SUPPORT, 0.0, stream=stream.cuda_stream)
            step(sets - 1)
            evals.append((_median_ms(step, torch, stream), dc, bpc, ilp, queue))
    svo.close()
    for k, v in saved.items():
        _capi.tune_set(k, v)
    best = choose(evals)
    return {"R": R, "dense_cells": best[1], "eval_blocks_per_cu": best[2], "eval_ilp": best[3],
            "eval_queue": best[4], "coarse_verts": best_lscg[1], "eval_ms": best[0], "eval_queries": n_queries,
            "eval_buffer_sets": sets, "lscg_ms": best_lscg[0], "candidates_eval": len(evals),
            "candidates_lscg": len(lscg), "workload": "torus-knot (3,5) tube, 200k triangles, jitter 1e-3 seed 7"}


def load(path=DEFAULT_TABLE):
    with open(path) as fh:
        return {int(k): v for k, v in json.load(fh)["depths"].items()}


def apply(depth, table=None):
    """Set the tunables recorded for `depth` (nearest recorded depth)."""
    from its_amd import _capi
    table = load() if table is None else table
    d = min(table, key=lambda k: (abs(k - depth), k))
    t = table[d]
    _capi.tune_set(_capi.TUNE_DENSE_CELLS, t["dense_cells"])
    _capi.tune_set(_capi.TUNE_EVAL_BLOCKS_PER_CU, t["eval_blocks_per_cu"])
    _capi.tune_set(_capi.TUNE_EVAL_ILP, t["eval_ilp"])
    _capi.tune_set(_capi.TUNE_EVAL_QUEUE, t["eval_queue"])
    _capi.tune_set(_capi.TUNE_COARSE_VERTS, t["coarse_verts"])
    return d


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--depths", type=int, nargs="+", default=[5, 6, 7, 8])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()
    from its_amd import _capi
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = _capi.Context(0)
    ctx.set_stream(stream.cuda_stream)
    table = {}
    for d in a.depths:
        table[d] = tune_depth(ctx, torch, stream, d)
        print(f"[autotune] depth {d}: {table[d]}", file=sys.stderr, flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    with open(a.out, "w") as fh:
        json.dump({"device": torch.cuda.get_device_name(0), "depths": table}, fh, indent=1)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
