"""The allocation-free batched evaluation (its_batch_eval_device_noalloc)
against its_batch_eval_device: bit-identical results on two calls in a row,
the second with fewer queries, and the size limits rejected on the host
before anything is launched.  No reference counterpart: the entry point is
for callers that cannot allocate device memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(ctx):
    from its_amd import _capi, meshes

    b = _capi.Batch(ctx)
    items = []
    for seed, R in ((3, 32), (4, 64), (5, 16)):
        V, F = meshes.noisy_icosphere(seed=seed, level=3)
        V = _capi.model_to_unit_cube(V)
        origin, end, width = _capi.uniform_bbox(V)
        g = _capi.Octree(ctx, meshes.triangles(V, F), origin, width, R)
        g.lscg(g.sdf_at_vertices(V, F), tol=1e-6)
        b.add(g)
        items.append((g, origin, end))
    b.fields()
    return b, items


def _queries(items, counts, seed):
    rng = np.random.default_rng(seed)
    Q = np.concatenate([rng.uniform(o - 0.05, e + 0.05, size=(n, 3)) for (_, o, e), n in zip(items, counts)])
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return np.ascontiguousarray(Q.astype(np.float32).T), offs


def test_noalloc_matches_batch_eval(ctx):
    import torch

    b, items = _batch(ctx)
    for call, counts in enumerate(((50_000, 70_000, 20_000), (3_000, 0, 1_234))):
        Q, offs = _queries(items, counts, seed=10 + call)
        print(f"[noalloc] call {call}: {offs[-1]} queries over {len(items)} meshes", flush=True)
        dq = torch.from_numpy(Q).cuda()
        ref = torch.full((int(offs[-1]),), -7.0, device="cuda")
        out = torch.full((int(offs[-1]),), -9.0, device="cuda")
        torch.cuda.synchronize()
        b.eval_device(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, ref.data_ptr(), 0.5)
        b.eval_device_noalloc(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, out.data_ptr(), 0.5)
        r, o = ref.cpu().numpy(), out.cpu().numpy()
        assert np.array_equal(r.view(np.uint32), o.view(np.uint32)), call
        assert np.all(o != -9.0)
    b.close()
    for g, _, _ in items:
        g.close()


def test_noalloc_limits(ctx):
    """Too many queries for the fixed work-item table: rejected with
    ITS_E_RANGE on the host, before any launch (the query pointers are never
    read, so small buffers stand in)."""
    import torch

    from its_amd import _capi

    b, items = _batch(ctx)
    n = _capi.NOALLOC_MAX_ITEMS * _capi.NOALLOC_CHUNK + 1
    offs = np.array([0, n, n, n], np.int64)
    dq = torch.zeros((3, 16), device="cuda")
    out = torch.zeros(16, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(_capi.ItsError) as e:
        b.eval_device_noalloc(dq[0].data_ptr(), dq[1].data_ptr(), dq[2].data_ptr(), offs, out.data_ptr(), 0.5)
    assert e.value.code == _capi.E_RANGE and "ITS_NOALLOC_MAX_ITEMS" in str(e.value)
    b.close()
    for g, _, _ in items:
        g.close()
