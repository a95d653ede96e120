"""Text output formats of the host classes (its_amd/host/its_io.hpp) against
the reference's writers restated in oracle/formats.py (CPU only).

The restatement is pinned first: the radical inverse against its closed
form, and the cube writer against the corner numbering of writeCube
(reference src/utils/IO.hpp:44-124).  A C++ driver then writes fixed inputs
through the library's writers and every file must equal the restated text
byte for byte.
"""
import os
import subprocess
from fractions import Fraction

import pytest

import formats as FMT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "its_amd", "host")
DRIVER = os.path.join(ROOT, "tests", "format_driver.cpp")


def _radical_inverse(i, b):
    """Exact value: the base-b digits of i mirrored about the radix point."""
    r, scale = Fraction(0), Fraction(1, b)
    while i:
        r += (i % b) * scale
        i //= b
        scale /= b
    return r


def test_halton_closed_form():
    assert FMT.halton(0, 2) == 0.0
    assert FMT.halton(1, 2) == 0.5 and FMT.halton(2, 2) == 0.25 and FMT.halton(3, 2) == 0.75
    assert FMT.halton(4, 3) == pytest.approx(4 / 9, abs=1e-16)
    assert FMT.halton(6, 5) == pytest.approx(6 / 25, abs=1e-16)
    for b in (2, 3, 5):
        for i in range(200):
            assert abs(FMT.halton(i, b) - float(_radical_inverse(i, b))) < 1e-15


def test_uniform_points_fill_scaled_box():
    lo, hi = [0.0, 0.0, 0.0], [1.0, 2.0, 4.0]
    P = FMT.uniform_points(lo, hi, 500, 2.0, [1.0, 0.0, -1.0])
    # the box scaled by 2 about its centre, then moved
    a, b = [-0.5 + 1.0, -1.0, -2.0 - 1.0], [1.5 + 1.0, 3.0, 6.0 - 1.0]
    assert P[0] == a   # index 0 is the lower corner
    for p in P:
        for k in range(3):
            assert a[k] <= p[k] < b[k]
    assert FMT.uniform_points(lo, hi, 10, 0.0, [0, 0, 0]) == []


def test_cube_text_structure():
    t = FMT.cube_obj_text([[1.0, 2.0, 3.0]], [0.5]).splitlines()
    v = [list(map(float, l.split()[1:])) for l in t if l.startswith("v ")]
    f = [list(map(int, l.split()[1:])) for l in t if l.startswith("f ")]
    assert len(v) == 8 and len(f) == 6
    assert v[3] == [1.0, 2.0, 3.0] and v[7] == [1.5, 2.5, 3.5]   # v4 = origin, v8 = far corner
    # every quad is a face of the cube: its 4 corners share one coordinate
    for q in f:
        pts = [v[i - 1] for i in q]
        assert any(len({p[k] for p in pts}) == 1 for k in range(3))
    assert sorted(i for q in f for i in q) == sorted(list(range(1, 9)) * 3)


def test_fixed9_alignment():
    s = FMT.fixed9_text([0.5, -12.25, 3.0]).splitlines()
    assert s == ["  0.500000000", "-12.250000000", "  3.000000000"]


@pytest.fixture(scope="module")
def driver_out(tmp_path_factory):
    d = tmp_path_factory.mktemp("formats")
    exe = str(d / "format_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", HOST, "-o", exe, DRIVER], check=True, timeout=120)
    subprocess.run([exe, str(d)], check=True, timeout=60)
    return d


def _read(d, name):
    with open(os.path.join(d, name)) as fh:
        return fh.read()


def test_values_and_fixed9(driver_out):
    vals = [0.0, -0.0, 1.0, 0.1, 1e-7, -123.456789, 1234567.0, 3.14159265358979, 1e20, -2.5e-5, 0.5000004,
            99999.95, 1e-300]
    assert _read(driver_out, "values.txt") == FMT.values_text(vals)
    assert _read(driver_out, "fixed9.txt") == FMT.fixed9_text([0.5, -12.25, 1e-10, 123456.789, -0.000000001, 0.0])
    assert _read(driver_out, "fixed9_empty.txt") == FMT.fixed9_text([])


def test_cubes_and_soup(driver_out):
    assert _read(driver_out, "cubes.obj") == FMT.cube_obj_text([[0.0, 0.0, 0.0], [-0.25, 0.125, 1.0 / 3.0]],
                                                                [1.0, 0.0625])
    soup = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.1, 0.2, 0.3], [1e-8, -2.5, 7], [3, 2, 1]]
    assert _read(driver_out, "soup.obj") == FMT.soup_obj_text([[float(x) for x in p] for p in soup])


def test_points_and_halton(driver_out):
    P = FMT.uniform_points([-0.1, 0.0, 0.2], [1.1, 0.9, 1.3], 40, 1.5, [0.01, -0.02, 0.0])
    assert _read(driver_out, "points.xyz") == FMT.point_cloud_text(P, ".xyz")
    assert _read(driver_out, "points.obj") == FMT.point_cloud_text(P, ".obj")
    assert _read(driver_out, "points.txt") == ""
    rows = [list(map(float, l.split())) for l in _read(driver_out, "halton.txt").splitlines()]
    assert rows == [[FMT.halton(i, 2), FMT.halton(i, 3), FMT.halton(i, 5)] for i in range(64)]
