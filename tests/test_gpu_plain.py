"""The whole plain method at its small size (noisy icosphere level 4, R = 64),
through the library and through the binary64 CPU restatement, compared
stage by stage (tests/plain_icosphere_report.py).

Bars: the octree bit for bit; the right-hand side within 1e-12; lambda no
further from the converged solution than 10x the restatement's own 1e-6
iterate; f at the mesh vertices within 1e-5 (binary32 evaluation of the
library's lambda against the binary64 evaluation of the restatement's); the
marching-cubes soup of the same lambda vertex for vertex within 1e-4 voxel.
"""
import pytest

import plain_icosphere_report as P

pytestmark = pytest.mark.gpu


def test_plain_icosphere_end_to_end(ctx):
    print("[plain] noisy icosphere level 4, seed 1, R=64", flush=True)
    rep = P.run(ctx, seed=1, R=64)
    print({k: rep[k] for k in ("n", "iterations", "b", "lambda_vs_converged", "f_vertices", "mc", "speedup")})
    assert rep["svo_bit_exact"]
    assert rep["n"] > 0 and rep["depth"] == 7
    assert rep["b"]["max_abs"] <= 1e-12
    assert rep["normal_rel"]["gpu"] < 1e-6
    assert rep["lambda_vs_converged"]["max_abs"] <= max(10 * rep["cpu_lambda_vs_converged"]["max_abs"], 1e-6)
    assert rep["f_vertices"]["max_abs"] <= 1e-5
    assert rep["f_vertices_same_lambda"]["max_abs"] <= 1e-5
    same = rep["mc_same_lambda"]
    assert same["soup_vertices"][0] == same["soup_vertices"][1] > 0
    assert same["vertex_max_in_voxels"] <= 1e-4
