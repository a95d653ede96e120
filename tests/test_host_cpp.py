"""C++ host classes (its_amd/host/its_host.hpp) over the C ABI.

The example its_amd/host/shell_demo.cpp is a main.cpp-style caller
(reference src/main.cpp:165-190): ThinShells(file, R, R, R, true, 1.0),
creatShell, multiPointQuery and MC::marching_cubes.  It is built by
__graft_entry__.build(); these tests run it and compare its summary with the
same steps driven through the Python binding.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "its_amd", "host", "shell_demo")
REF_CALLER = os.path.join(ROOT, "its_amd", "host", "ref_caller")
GOLDEN = os.path.join(ROOT, "tests", "golden", "bunny.npz")


def _write_off(path, V, F):
    with open(path, "w") as fh:
        fh.write(f"OFF\n{len(V)} {len(F)} 0\n")
        for v in V:
            fh.write("%.17g %.17g %.17g\n" % tuple(v))
        for f in F:
            fh.write("3 %d %d %d\n" % tuple(f))


def _queries(origin, width, nq):
    i = np.arange(1, nq + 1, dtype=np.float64)[:, None]
    a = np.arange(3, dtype=np.float64)[None, :]
    u = np.fmod(0.6180339887498949 * i * (a + 1) + 0.1 * a, 1.0)
    return origin - 0.05 * width + 1.1 * u * width


def test_demo_is_built():
    assert os.path.exists(DEMO), "run __graft_entry__.build()"
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_demo_reports_missing_file(tmp_path):
    r = subprocess.run([DEMO, str(tmp_path / "none.off"), "16"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "its error" in r.stderr


@pytest.mark.gpu
def test_demo_matches_python_path(ctx, tmp_path):
    from its_amd import _capi, meshes
    d = np.load(GOLDEN)
    V, F = d["V"], d["F"].astype(np.int32)
    off = tmp_path / "bunny.off"
    _write_off(off, V, F)
    R, nq = 64, 3000
    print(f"[host demo] bunny R={R}", flush=True)
    r = subprocess.run([DEMO, str(off), str(R), str(nq)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])

    Vn = _capi.model_to_unit_cube(V)
    origin, end, width = _capi.uniform_bbox(Vn)
    svo = _capi.Octree(ctx, meshes.triangles(Vn, F), origin, width, R)
    b = svo.sdf_at_vertices(Vn, F)
    lam, st = svo.lscg(b, tol=1e-6)
    inner, outer, _ = svo.shell_iso(Vn, F)
    q = _queries(origin, width, nq)
    fq = svo.eval(q, end, _capi.EVAL_SUPPORT, np.float32(outer))
    cls = _capi.classify(fq, inner, outer)
    tri, mst = svo.marching_cubes(origin, width, (R, R, R), 0.0)
    assert got["tree_depth"] == svo.depth and got["num_nodes"] == svo.num_nodes
    assert got["num_verts"] == svo.num_verts and got["num_voxels"] == svo.num_voxels
    assert got["lambda_sum"] == float(np.sum(lam)) or abs(got["lambda_sum"] - np.sum(lam)) <= 1e-12 * np.abs(lam).sum()
    assert abs(got["lambda_abs"] - np.abs(lam).sum()) <= 1e-12 * np.abs(lam).sum()
    assert np.float32(got["inner_iso"]) == np.float32(inner) and np.float32(got["outer_iso"]) == np.float32(outer)
    assert [got["inside"], got["shell"], got["outside"]] == [int((cls == -1).sum()), int((cls == 0).sum()),
                                                              int((cls == 1).sum())]
    assert got["mc_verts"] == mst["num_verts"]
    svo.close()


def _read(p):
    with open(p) as fh:
        return fh.read()


@pytest.mark.gpu
def test_demo_dumps_and_validation(ctx, tmp_path):
    """The demo's text outputs (SDFValue, Coefficients, latent points, f at the
    latent points, octree cubes per level, MC soup) and its point in/out
    validation (main.cpp:44-134) against the same steps through the Python
    binding, with the formats of oracle/formats.py."""
    import formats as FMT
    from its_amd import _capi, meshes
    d = np.load(GOLDEN)
    V, F = d["V"], d["F"].astype(np.int32)
    off = tmp_path / "bunny.off"
    _write_off(off, V, F)
    R, nq = 32, 4000
    dump, obj = tmp_path / "dump", tmp_path / "iso.obj"
    print(f"[host demo dumps] bunny R={R}", flush=True)
    r = subprocess.run([DEMO, str(off), str(R), str(nq), str(obj), str(dump)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])

    Vn = _capi.model_to_unit_cube(V)
    origin, end, width = _capi.uniform_bbox(Vn)
    svo = _capi.Octree(ctx, meshes.triangles(Vn, F), origin, width, R)
    b = svo.sdf_at_vertices(Vn, F)
    lam, st = svo.lscg(b, tol=1e-6)
    assert _read(dump / "SDFValue.txt") == FMT.values_text(b)
    assert _read(dump / "Coefficients.txt") == FMT.values_text(lam)

    # latent points: the model vertices, then the intersection points (edge
    # points without exact duplicates, face points)
    P, ne = svo.shell_points(Vn, F)
    edge = np.unique(P[:ne], axis=0)
    lines = _read(dump / "latent_point.xyz").splitlines()
    assert len(lines) == len(Vn) + got["latent_points"]
    assert len(edge) + len(P) - ne <= got["latent_points"] <= len(P)
    assert lines[: len(Vn)] == FMT.point_cloud_text(Vn, ".xyz").splitlines()
    assert set(lines[len(Vn):]) == set(FMT.point_cloud_text(np.concatenate([edge, P[ne:]]), ".xyz").splitlines())
    bs = _read(dump / "BSplineValue.txt").splitlines()
    assert len(bs) == len(lines) and len({len(l) for l in bs}) == 1   # aligned column
    fV = svo.eval(Vn, end, _capi.EVAL_SUPPORT, 0.0).astype(np.float64)
    assert [l.strip() for l in bs[: len(Vn)]] == ["%.9f" % v for v in fV]

    nodes = svo.nodes()
    first = 0
    for lvl, cnt in enumerate(svo.depth_num_nodes()):
        o, w = nodes["origin"][first:first + cnt], nodes["width"][first:first + cnt]
        assert _read(dump / "svo" / f"svo_{lvl + 1}.obj") == FMT.cube_obj_text(o.tolist(), w.tolist())
        first += cnt
    tri, mst = svo.marching_cubes(origin, width, (R, R, R), 0.0)
    assert _read(obj) == FMT.soup_obj_text(tri.astype(np.float64).tolist())

    Q = np.array(FMT.uniform_points(list(origin), list(end), nq, 1.0, [0.0, 0.0, 0.0]))
    assert _read(dump / "query.xyz") == FMT.point_cloud_text(Q, ".xyz")
    inner, outer, _ = svo.shell_iso(Vn, F)
    cls = _capi.classify(svo.eval(Q, end, _capi.EVAL_SUPPORT, np.float32(outer)), inner, outer)
    exact = np.sign(ctx.signed_distance(Vn, F, Q)).astype(int)
    correct = int((cls == exact).sum())
    assert got["check_correct"] == correct and got["check_errors"] == nq - correct
    assert got["check_rate"] == pytest.approx(100.0 * correct / nq, abs=1e-6)
    assert len(_read(dump / "query_err.xyz").splitlines()) == nq - correct
    assert correct > nq // 2
    svo.close()


def _read_obj_vertices(path):
    with open(path) as fh:
        return np.array([[float(x) for x in l.split()[1:]] for l in fh if l.startswith("v ")])


@pytest.mark.gpu
def test_reference_shaped_caller(ctx, tmp_path):
    """its_amd/host/ref_caller.cpp calls the host classes with the
    reference's argument lists (ThinShells(file, V3i, is2UnitCube, scale),
    mcVisualization, MC::marching_cubes(nodeVertexArray, numNodes,
    nodeWidthArray, numNodeVerts, lambda, double3, double3, uint3, iso, file),
    multiPointQuery(out_file, points)).  Its files are checked against the
    same steps through the Python binding."""
    from its_amd import _capi, meshes
    d = np.load(GOLDEN)
    V, F = d["V"], d["F"].astype(np.int32)
    off = tmp_path / "bunny.off"
    _write_off(off, V, F)
    R, res = 32, 40
    out = tmp_path / "out"
    print(f"[ref caller] bunny R={R} res={res}", flush=True)
    r = subprocess.run([REF_CALLER, str(off), str(R), str(res), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    Vn = _capi.model_to_unit_cube(V)
    origin, end, width = _capi.uniform_bbox(Vn)
    svo = _capi.Octree(ctx, meshes.triangles(Vn, F), origin, width, R)
    b = svo.sdf_at_vertices(Vn, F)
    lam, _ = svo.lscg(b, tol=1e-6)
    inner, outer, _ = svo.shell_iso(Vn, F)
    assert got["num_verts"] == svo.num_verts and got["tree_depth"] == svo.depth
    assert got["inner_iso"] == inner and got["outer_iso"] == outer
    soups, _ = svo.marching_cubes_multi(origin, width, (res, res, res), [inner, outer, 0.0])
    for name, soup in zip(["mc_innerShell.obj", "mc_outerShell.obj", "mc_isosurface.obj"], soups):
        P = _read_obj_vertices(out / name)
        assert P.shape == soup.shape and len(P) > 0, name
        assert np.abs(P - soup).max() <= 1e-5, name          # text at 6 significant digits
    # the reference-signature overload: the same soup from the host arrays
    P = _read_obj_vertices(out / "mc_reference_signature.obj")
    assert P.shape == soups[2].shape and np.abs(P - soups[2]).max() <= 1e-5
    # query point cloud, coloured by inner < f < outer over all basis functions
    Q = _read_obj_vertices(out / "query_result.obj")
    assert Q.shape == (200, 6)
    i = np.arange(1, 201, dtype=np.float64)[:, None]
    a = np.arange(3, dtype=np.float64)[None, :]
    pts = origin + width * np.fmod(0.7548776662466927 * i * (a + 1) + 0.2 * a, 1.0)   # the caller's points
    assert np.abs(Q[:, :3] - pts).max() <= 1e-5
    f = svo.eval(pts, end, _capi.EVAL_RAW, 0.0)
    on = (inner < f) & (f < outer)
    np.testing.assert_allclose(Q[on, 3:], np.tile([0.56471, 0.93333, 0.56471], (on.sum(), 1)))
    np.testing.assert_allclose(Q[~on, 3:], np.tile([1, 0.27059, 0], ((~on).sum(), 1)))
    assert _read_obj_vertices(out / "single_query.obj").shape == (1, 6)
    svo.close()
