"""ctypes binding of the FP64 oracle (libits_oracle.so).

TEST INFRASTRUCTURE ONLY.  It is imported by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
(its_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libits_oracle.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "ora_morton_encode": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "ora_morton_tables": (None, [_u32p, _u32p, _u32p, _u8p, _u8p, _u8p]),
            "ora_neighbor_luts": (None, [_i16p, _i16p, _i16p, _i16p]),
            "ora_model_to_unit_cube": (None, [_f64p, C.c_int64, C.c_double]),
            "ora_uniform_bbox": (None, [_f64p, C.c_int64, _f64p, _f64p, _f64p]),
            "ora_voxelize": (C.c_int64, [_f64p, C.c_int64, _f64p, _f64p, C.c_int, _u32p, C.c_int64]),
            "ora_svo_build": (vp, [_f64p, C.c_int64, _f64p, _f64p, C.c_int]),
            "ora_svo_free": (None, [vp]),
            "ora_svo_depth": (C.c_int, [vp]),
            "ora_svo_num_nodes": (C.c_int64, [vp]),
            "ora_svo_num_verts": (C.c_int64, [vp]),
            "ora_svo_num_voxels": (C.c_int64, [vp]),
            "ora_svo_level_info": (None, [vp, _i64p, _i64p]),
            "ora_svo_nodes": (None, [vp, _u32p, _u8p, _f64p, _f64p, _u32p, _u32p, _u32p]),
            "ora_svo_verts": (None, [vp, _f64p, _u32p]),
            "ora_base_function": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "ora_assemble": (vp, [vp]),
            "ora_mat_free": (None, [vp]),
            "ora_mat_nnz": (C.c_int64, [vp]),
            "ora_mat_n": (C.c_int64, [vp]),
            "ora_mat_csr": (None, [vp, _i64p, _i32p, _f64p]),
            "ora_mat_mul": (None, [vp, _f64p, _f64p]),
            "ora_mat_tmul": (None, [vp, _f64p, _f64p]),
            "ora_lscg": (C.c_int, [vp, _f64p, _f64p, C.c_double, C.c_int64, _f64p]),
            "ora_eval_support": (None, [vp, _f64p, _f64p, C.c_int64, _f64p, C.c_double, _f64p]),
            "ora_eval_latent": (C.c_int, [vp, _f64p, _f64p, C.c_int64, _f64p]),
            "ora_eval_dense": (None, [vp, _f64p, _f64p, C.c_int64, _f64p]),
            "ora_eval_dense_fast": (None, [vp, _f64p, _f64p, C.c_int64, _f64p]),
            "ora_marching_cubes": (vp, [vp, _f64p, _f64p, _f64p, _i32p, C.c_double]),
            "ora_marching_cubes_slab": (vp, [vp, _f64p, _f64p, _f64p, _i32p, C.c_double, C.c_int, C.c_int]),
            "ora_mc_free": (None, [vp]),
            "ora_mc_num_verts": (C.c_int64, [vp]),
            "ora_mc_num_valid_voxels": (C.c_int64, [vp]),
            "ora_mc_soup": (None, [vp, _f64p]),
            "ora_mc_cube_index": (None, [vp, _u8p]),
            "ora_mc_tri_table": (C.POINTER(C.c_int), []),
            "ora_mc_nverts_table": (C.POINTER(C.c_int), []),
            "ora_signed_distance": (None, [_f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.c_int64, _f64p]),
            "ora_winding_number": (C.c_double, [_f64p, _i32p, C.c_int64, _f64p]),
            "ora_num_threads": (C.c_int, []),
            "ora_intersections": (C.c_int64, [vp, _f64p, C.c_int64, _i32p, C.c_int64, vp, C.c_int64, _i64p]),
            "ora_fine_edges": (C.c_int64, [vp, vp, vp, C.c_int64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def normalize(V, scale_factor=1.0):
    """model2UnitCube (BaseModel.cpp:84-92) on a copy of V."""
    V = _f64(V).copy()
    lib().ora_model_to_unit_cube(V.reshape(-1), len(V), scale_factor)
    return V


def uniform_bbox(V):
    """setUniformBoundingBox (BaseModel.cpp:24-51) -> (origin, end, width)."""
    o, e, w = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().ora_uniform_bbox(_f64(V).reshape(-1), len(V), o, e, w)
    return o, e, w


def voxelize(tris, origin, width, R):
    tris = _f64(tris)
    cap = 1 << 20
    while True:
        out = np.zeros(cap, dtype=np.uint32)
        n = lib().ora_voxelize(tris.reshape(-1), len(tris), _f64(origin), _f64(width), R, out, cap)
        if n >= 0:
            return out[:n]
        cap = -n


class SVO:
    """Oracle octree: node and vertex arrays as numpy arrays."""

    def __init__(self, tris, origin, width, R):
        tris = _f64(tris)
        self.R = R
        self.origin = _f64(origin)
        self.width = _f64(width)
        self._h = lib().ora_svo_build(tris.reshape(-1), len(tris), self.origin, self.width, R)
        L = lib()
        self.depth = L.ora_svo_depth(self._h)
        self.num_nodes = L.ora_svo_num_nodes(self._h)
        self.num_verts = L.ora_svo_num_verts(self._h)
        self.num_voxels = L.ora_svo_num_voxels(self._h)
        self.depth_num_nodes = np.zeros(max(self.depth, 1), dtype=np.int64)
        self.esum_verts = np.zeros(self.depth + 1, dtype=np.int64)
        L.ora_svo_level_info(self._h, self.depth_num_nodes, self.esum_verts)
        n = max(self.num_nodes, 1)
        self.morton = np.zeros(n, np.uint32)
        self.is_leaf = np.zeros(n, np.uint8)
        self.node_origin = np.zeros((n, 3))
        self.node_width = np.zeros(n)
        self.parent = np.zeros(n, np.uint32)
        self.childs = np.zeros((n, 8), np.uint32)
        self.neighbors = np.zeros((n, 27), np.uint32)
        L.ora_svo_nodes(self._h, self.morton, self.is_leaf, self.node_origin.reshape(-1), self.node_width,
                        self.parent, self.childs.reshape(-1), self.neighbors.reshape(-1))
        m = max(self.num_verts, 1)
        self.vert_xyz = np.zeros((m, 3))
        self.vert_owner = np.zeros(m, np.uint32)
        L.ora_svo_verts(self._h, self.vert_xyz.reshape(-1), self.vert_owner)
        self.num_nodes_arr = self.num_nodes

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ora_svo_free(self._h)
            self._h = None

    def eval_support(self, lam, pts, box_end, out_val):
        pts = _f64(pts)
        f = np.zeros(len(pts))
        lib().ora_eval_support(self._h, _f64(lam), pts.reshape(-1), len(pts), _f64(box_end), out_val, f)
        return f

    def eval_latent(self, lam, pts):
        pts = _f64(pts)
        f = np.zeros(len(pts))
        rc = lib().ora_eval_latent(self._h, _f64(lam), pts.reshape(-1), len(pts), f)
        return f, rc

    def eval_dense(self, lam, pts):
        pts = _f64(pts)
        f = np.zeros(len(pts))
        lib().ora_eval_dense(self._h, _f64(lam), pts.reshape(-1), len(pts), f)
        return f

    def eval_dense_fast(self, lam, pts):
        pts = _f64(pts)
        f = np.zeros(len(pts))
        lib().ora_eval_dense_fast(self._h, _f64(lam), pts.reshape(-1), len(pts), f)
        return f


class Matrix:
    """Oracle collocation matrix A_ij = B_j(v_i) (ThinShells.cpp:511-605)."""

    def __init__(self, svo: SVO):
        self._svo = svo
        self._h = lib().ora_assemble(svo._h)
        self.n = lib().ora_mat_n(self._h)
        self.nnz = lib().ora_mat_nnz(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ora_mat_free(self._h)
            self._h = None

    def csr(self):
        rp = np.zeros(self.n + 1, np.int64)
        c = np.zeros(max(self.nnz, 1), np.int32)
        v = np.zeros(max(self.nnz, 1))
        lib().ora_mat_csr(self._h, rp, c, v)
        return rp, c[: self.nnz], v[: self.nnz]

    def mul(self, x):
        y = np.zeros(self.n)
        lib().ora_mat_mul(self._h, _f64(x), y)
        return y

    def tmul(self, x):
        y = np.zeros(self.n)
        lib().ora_mat_tmul(self._h, _f64(x), y)
        return y

    def lscg(self, b, x0=None, tol=1e-6, max_iter=None):
        x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
        st = np.zeros(4)
        lib().ora_lscg(self._h, _f64(b), x, tol, 2 * self.n if max_iter is None else max_iter, st)
        stats = {"iterations": int(st.view(np.int64)[0]), "normal_rel": st[1], "residual_rel": st[2],
                 "residual_abs": st[3]}
        return x, stats


def intersections(svo: SVO, V, F):
    """Shell iso-band points (ThinShells::cpIntersectionPoints): unique
    mesh-edge x node-face points, then node-edge x mesh-face points."""
    V = _f64(V)
    F = np.ascontiguousarray(F, np.int32)
    ne = np.zeros(1, np.int64)
    n = lib().ora_intersections(svo._h, V.reshape(-1), len(V), F.reshape(-1), len(F), None, 0, ne)
    out = np.zeros((max(n, 1), 3))
    lib().ora_intersections(svo._h, V.reshape(-1), len(V), F.reshape(-1), len(F), out.ctypes.data_as(C.c_void_p), n,
                            ne)
    return out[:n], int(ne[0])


def shell_iso(svo: SVO, lam, V, F):
    """(inner, outer) iso values: min and max of the latent values at the
    mesh vertices and the intersection points (cpLatentBSplineValue,
    src/ThinShells.cpp:655-755, 743-744)."""
    Q, _ = intersections(svo, V, F)
    f, rc = svo.eval_latent(lam, np.concatenate([_f64(V), Q]))
    if rc != 0:
        raise ValueError("a latent point's fine voxel is not an octree node")
    return float(f.min()), float(f.max())


def classify(f, inner, outer):
    """{-1, 0, 1} classes of query values (multiPointQuery,
    src/ThinShells.cpp:1412-1417): 1 at or above the outer iso value, -1 at
    or below the inner one, 0 in between (a NaN compares false: 0)."""
    f = np.asarray(f, np.float64)
    return np.where(f >= outer, 1, np.where(f <= inner, -1, 0)).astype(np.int8)


def fine_edges(svo: SVO):
    """Level-0 node edges (fineNodeEdgeArray, src/cuAcc/SVO.cu:965-1110):
    (n, 2, 3) end points and the owner node of each, in the reference's order."""
    n = lib().ora_fine_edges(svo._h, None, None, 0)
    ends = np.zeros((max(n, 1), 2, 3))
    owner = np.zeros(max(n, 1), np.uint32)
    lib().ora_fine_edges(svo._h, ends.ctypes.data_as(C.c_void_p), owner.ctypes.data_as(C.c_void_p), n)
    return ends[:n], owner[:n]


def lscg_operator(mul, tmul, invdiag, b, tol=1e-6, max_iter=None, x0=None):
    """Eigen's LeastSquaresConjugateGradient (the algorithm of ora_lscg, as
    called at ThinShells.cpp:615-617) in binary64 numpy over an operator:
    mul(x) = A x, tmul(r) = A^T r, invdiag = 1/||A_:j||^2 (1 for an empty
    column).  Returns (x, stats) like Matrix.lscg."""
    b = np.asarray(b, np.float64)
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    max_iter = 2 * n if max_iter is None else max_iter
    residual = b - mul(x)
    nres = tmul(residual)
    rhs2 = float(np.dot(tmul(b), tmul(b)))
    it = 0
    if rhs2 == 0:
        x[:] = 0
    else:
        thr = tol * tol * rhs2
        rn2 = float(np.dot(nres, nres))
        if rn2 >= thr:
            p = invdiag * nres
            abs_new = float(np.dot(nres, p))
            while it < max_iter:
                t = mul(p)
                alpha = abs_new / float(np.dot(t, t))
                x += alpha * p
                residual -= alpha * t
                nres = tmul(residual)
                rn2 = float(np.dot(nres, nres))
                if rn2 < thr:
                    break
                z = invdiag * nres
                abs_old = abs_new
                abs_new = float(np.dot(nres, z))
                p = z + (abs_new / abs_old) * p
                it += 1
    r = mul(x) - b
    nr = tmul(r)
    return x, {"iterations": it, "normal_rel": float(np.sqrt(np.dot(nr, nr) / rhs2)) if rhs2 > 0 else 0.0,
               "residual_rel": float(np.linalg.norm(r) / max(np.linalg.norm(b), 1e-300)),
               "residual_abs": float(np.linalg.norm(r))}


def marching_cubes(svo: SVO, lam, origin, width, res, iso, zslab=None):
    """MC::marching_cubes restated: (soup, cube indices, valid voxels).  With
    zslab=(z0, z1) only the voxels of that z-range of the same grid: their
    cube indices and the contiguous piece of the whole soup they emit."""
    L = lib()
    res = np.asarray(res, np.int32)
    if zslab is None:
        h = L.ora_marching_cubes(svo._h, _f64(lam), _f64(origin), _f64(width), res, iso)
        nvox = int(np.prod(res))
    else:
        z0, z1 = int(zslab[0]), int(zslab[1])
        h = L.ora_marching_cubes_slab(svo._h, _f64(lam), _f64(origin), _f64(width), res, iso, z0, z1)
        nvox = int(res[0]) * int(res[1]) * (z1 - z0)
    try:
        nv = L.ora_mc_num_verts(h)
        soup = np.zeros((max(nv, 1), 3))
        L.ora_mc_soup(h, soup.reshape(-1))
        cube = np.zeros(max(nvox, 1), np.uint8)
        L.ora_mc_cube_index(h, cube)
        return soup[:nv], cube[:nvox], L.ora_mc_num_valid_voxels(h)
    finally:
        L.ora_mc_free(h)


def tri_table():
    p = lib().ora_mc_tri_table()
    return np.array([p[i] for i in range(256 * 16)], dtype=np.int32).reshape(256, 16)


def nverts_table():
    p = lib().ora_mc_nverts_table()
    return np.array([p[i] for i in range(256)], dtype=np.int32)


def signed_distance(V, F, pts):
    V = _f64(V)
    F = np.ascontiguousarray(F, dtype=np.int32)
    pts = _f64(pts)
    out = np.zeros(len(pts))
    lib().ora_signed_distance(V.reshape(-1), len(V), F.reshape(-1), len(F), pts.reshape(-1), len(pts), out)
    return out


def winding_number(V, F, p):
    return lib().ora_winding_number(_f64(V).reshape(-1), np.ascontiguousarray(F, dtype=np.int32).reshape(-1), len(F),
                                    _f64(p))


def morton_encode(x, y, z):
    return lib().ora_morton_encode(x, y, z)


def morton_tables():
    ex, ey, ez = (np.zeros(256, np.uint32) for _ in range(3))
    dx, dy, dz = (np.zeros(512, np.uint8) for _ in range(3))
    lib().ora_morton_tables(ex, ey, ez, dx, dy, dz)
    return ex, ey, ez, dx, dy, dz


def neighbor_luts():
    lp = np.zeros(8 * 27, np.int16)
    lc = np.zeros(8 * 27, np.int16)
    vs = np.zeros(64, np.int16)
    es = np.zeros(48, np.int16)
    lib().ora_neighbor_luts(lp, lc, vs, es)
    return lp.reshape(8, 27), lc.reshape(8, 27), vs, es
