"""Two host threads, one context each (the ABI's convention: one context
per host thread, stream-ordered), run the whole path on different meshes at
the same time.  Octree, signed distance, fit (fixed-order reductions),
evaluation and marching cubes are deterministic, so every result must equal
the one-thread run bit for bit."""
import threading

import numpy as np
import pytest

from helpers import case

pytestmark = pytest.mark.gpu

CASES = [("icosphere", 32), ("torus", 32), ("bunny", 32), ("random", 32)]


def _run(c_ctx, name, R):
    from its_amd import _capi
    c = case(name, R)
    g = _capi.Octree(c_ctx, c.tris, c.origin, c.width, R)
    b = g.sdf_at_vertices(c.V, c.F)
    lam, st = g.lscg(b, tol=1e-6)
    q = np.random.default_rng(5).uniform(-0.05, 1.05, size=(20000, 3)).astype(np.float32)
    f = g.eval(q, c.end)
    tri, _ = g.marching_cubes(c.origin, c.width, (R, R, R), 0.0)
    g.close()
    return b, lam, f, tri


def test_two_host_threads(ctx):
    from its_amd import _capi
    print("[threads] one-thread reference", flush=True)
    ref = {k: _run(ctx, *k) for k in CASES}
    ctxs = [_capi.Context(0) for _ in range(2)]
    results, errors = {}, []

    def worker(t):
        try:
            for _ in range(2):
                for k in CASES[t::2]:
                    results[k] = _run(ctxs[t], *k)
        except Exception as e:   # re-raised below
            errors.append(e)

    print("[threads] two threads", flush=True)
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads)
    for c in ctxs:
        c.close()
    if errors:
        raise errors[0]
    for k in CASES:
        for a, r in zip(results[k], ref[k]):
            np.testing.assert_array_equal(a, r)
