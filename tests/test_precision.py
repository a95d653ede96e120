"""Precision of the mixed-precision path at small size (GPU), and the grid
marching cubes used by the precision report (CPU).

The full study (R = 512, about 1.5 M triangles) is tests/precision_report.py;
here the same comparisons run on a small blob union with stated bounds:
lambda within 1e-4 relative of the binary64 fit at the same tolerance, f at
the vertices within 1e-5, marching-cubes vertices of cubes with the same case
within 1e-5 of a voxel.
"""
import numpy as np
import pytest

from its_amd import grid_mc


def test_grid_mc_sphere_is_closed():
    res = 40
    g = np.arange(res + 1) / res
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    c = np.array([0.5, 0.52, 0.47])
    sdf = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - 0.3
    V, F = grid_mc.marching_cubes(sdf, (0, 0, 0), 1 / res, 0.0, device="cpu")
    assert np.abs(np.linalg.norm(V - c, axis=1) - 0.3).max() < 1e-3
    E = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    u, cnt = np.unique(E, axis=0, return_counts=True)
    assert np.all(cnt == 2)                      # watertight
    assert len(V) - len(u) + len(F) == 2         # sphere: Euler characteristic 2
    soup, case = grid_mc.marching_cubes(sdf, (0, 0, 0), 1 / res, 0.0, weld=False, device="cpu")
    assert len(soup) == 3 * len(F)
    nv = (grid_mc.tri_table() >= 0).sum(1)
    assert nv[case].sum() == len(soup)


def test_grid_mc_table_is_the_kernel_table():
    import oracle as O
    t = grid_mc.tri_table()
    assert np.array_equal(t, np.asarray(O.tri_table(), np.int64).reshape(256, 16))


@pytest.mark.gpu
def test_precision_small(ctx):
    import precision_report as P
    rep = P.run(ctx, R=64, target_tris=2e4, seed=7)
    print(rep)
    assert rep["lambda_vs_f64_same_tol"]["rel_max"] <= 1e-4
    assert rep["f_vertices_vs_f64_same_tol"]["max_abs"] <= 1e-5
    assert rep["mc"]["library_soup_equals_recomputed"]
    assert rep["mc"]["vertex_max_in_voxels"] <= 1e-4
    assert rep["mc"]["cubes_with_other_case"] <= 1e-3 * 64 ** 3
