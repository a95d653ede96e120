"""Test infrastructure: the reference's text output formats restated in Python.

Only tests/ import this module; it checks the library's C++ writers
(its_amd/host/its_io.hpp) byte for byte.

- ``values_text``: one value per line, C++ ``ostream << double`` with default
  flags, i.e. printf "%g" (ThinShells::saveSDFValue / saveCoefficients,
  reference src/ThinShells.cpp:1017-1046).
- ``fixed9_text``: Eigen's column-vector output under ``fixed`` and
  ``setprecision(9)``: every coefficient right-aligned to the widest one,
  rows joined by newlines, then ``endl`` (saveBSplineValue,
  src/ThinShells.cpp:1068-1081).
- ``point_cloud_text``: "v x y z" for .obj, "x y z" for .xyz
  (src/utils/IO.hpp:7-23, 126-154).
- ``cube_obj_text``: writeCube with CUBE_WRITE (src/utils/IO.hpp:44-124; the
  x64 builds define CUBE_WRITE, 3DThinShell.vcxproj:159,185), one file per
  octree level in saveSVO (src/cuAcc/SVO.cu:1266-1318).
- ``soup_obj_text``: MC::writeToOBJFile (src/cuAcc/MarchingCubes/
  MarchingCubes.cu:635-661).
- ``halton`` / ``uniform_points``: getUniformRandomMatrix's Halton points
  (src/utils/common.hpp:320-370) in the box of AABox::scaleAndTranslate
  (src/utils/Geometry.hpp:18-30).
"""


def g(x):
    return "%g" % x


def values_text(values):
    return "".join(g(v) + "\n" for v in values)


def fixed9_text(values):
    s = ["%.9f" % v for v in values]
    w = max((len(t) for t in s), default=0)
    return "\n".join(t.rjust(w) for t in s) + "\n"


def point_cloud_text(points, ext):
    if ext == ".obj":
        return "".join("v %s %s %s\n" % (g(p[0]), g(p[1]), g(p[2])) for p in points)
    if ext == ".xyz":
        return "".join("%s %s %s\n" % (g(p[0]), g(p[1]), g(p[2])) for p in points)
    return ""


# writeCube's corners v1..v8 as offsets and its 6 quads (CUBE_WRITE)
_CUBE_OFFSETS = [(0, 1, 1), (0, 0, 1), (0, 1, 0), (0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1)]
_CUBE_QUADS = [(3, 4, 2, 1), (6, 5, 4, 3), (1, 8, 6, 3), (1, 2, 7, 8), (4, 5, 7, 2), (8, 7, 5, 6)]


def cube_obj_text(origins, widths):
    out = []
    base = 0
    for o, w in zip(origins, widths):
        for d in _CUBE_OFFSETS:
            v = [o[k] + (w if d[k] else 0.0) for k in range(3)]
            out.append("v %s %s %s\n" % (g(v[0]), g(v[1]), g(v[2])))
        for q in _CUBE_QUADS:
            out.append("f %d %d %d %d\n" % tuple(base + i for i in q))
        base += 8
    return "".join(out)


def soup_obj_text(points):
    out = []
    for i in range(0, len(points) - 2, 3):
        for p in points[i:i + 3]:
            out.append("v %s %s %s\n" % (g(p[0]), g(p[1]), g(p[2])))
        out.append("f %d %d %d\n" % (i + 1, i + 2, i + 3))
    return "".join(out)


def halton(index, base):
    result, f, i = 0.0, 1.0 / base, index
    while i > 0:
        result += f * (i % base)
        i //= base
        f /= base
    return result


def uniform_points(lo, hi, num, scale, t):
    if scale <= 0.0:
        return []
    a, b = [], []
    for k in range(3):
        c = (lo[k] + hi[k]) / 2.0
        a.append((lo[k] - c) * scale + c + t[k])
        b.append((hi[k] - c) * scale + c + t[k])
    return [[a[k] + halton(i, (2, 3, 5)[k]) * (b[k] - a[k]) for k in range(3)] for i in range(num)]
