"""ctypes binding of the C ABI in include/its_c.h (libits_hip.so).

This is plumbing only: every computation runs in the HIP library.  If the
library is missing, importing works but every call raises
:class:`LibraryMissing`.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libits_hip.so")

ITS_OK = 0
ITS_NONE = 0xFFFFFFFF
ITS_MAX_LEVELS = 12
EVAL_SUPPORT, EVAL_DENSE, EVAL_RAW = 0, 1, 2
TUNE_DENSE_CELLS, TUNE_EVAL_BLOCKS_PER_CU, TUNE_COARSE_VERTS, TUNE_EVAL_ILP, TUNE_EVAL_QUEUE = 0, 1, 2, 3, 4
TUNE_BATCH_MAX_VERTS = 5
TUNE_SDF_RAY = 6
NOALLOC_MAX_MESHES, NOALLOC_MAX_ITEMS, NOALLOC_CHUNK = 1024, 65536, 16384
E_INVALID, E_HIP, E_NOMEM, E_STATE, E_RANGE = -1, -2, -3, -4, -5


class LibraryMissing(RuntimeError):
    pass


class ItsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"its error {code}: {msg}")
        self.code = code


class SvoInfo(C.Structure):
    _fields_ = [
        ("tree_depth", C.c_int),
        ("R", C.c_int),
        ("num_voxels", C.c_int64),
        ("num_nodes", C.c_int64),
        ("num_verts", C.c_int64),
        ("depth_num_nodes", C.c_int64 * ITS_MAX_LEVELS),
        ("esum_nodes", C.c_int64 * (ITS_MAX_LEVELS + 1)),
        ("esum_verts", C.c_int64 * (ITS_MAX_LEVELS + 1)),
        ("grid_origin", C.c_double * 3),
        ("voxel_width", C.c_double),
    ]


class LscgStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("refinements", C.c_int64),
        ("normal_rel", C.c_double),
        ("residual_rel", C.c_double),
        ("residual_abs", C.c_double),
        ("initial_normal_rel", C.c_double),
        ("seconds", C.c_double),
        ("converged", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LscgStructure(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("parent_entries", C.c_int64),
        ("levels", C.c_int),
        ("block_levels", C.c_int),
        ("bytes_assembled_per_iteration", C.c_double),
        ("bytes_matrix_free_per_iteration", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class McStats(C.Structure):
    _fields_ = [
        ("num_verts", C.c_int64),
        ("num_valid_voxels", C.c_int64),
        ("seconds_eval", C.c_double),
        ("seconds_extract", C.c_double),
        ("seconds_eval_device", C.c_double),
        ("seconds_extract_device", C.c_double),
    ]


_lib = None
vp = C.c_void_p
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")

# name -> argtypes (all return int status unless listed in _RET)
SIGNATURES = {
    "its_ctx_create": [C.c_int, C.POINTER(vp)],
    "its_ctx_destroy": [vp],
    "its_ctx_set_stream": [vp, vp],
    "its_ctx_synchronize": [vp],
    "its_device_count": [C.POINTER(C.c_int)],
    "its_model_to_unit_cube": [_f64p, C.c_int64, C.c_double],
    "its_uniform_bbox": [_f64p, C.c_int64, _f64p, _f64p, _f64p],
    "its_svo_build": [vp, _f64p, C.c_int64, _f64p, _f64p, C.c_int, C.POINTER(vp)],
    "its_svo_build_device": [vp, vp, C.c_int64, _f64p, _f64p, C.c_int, C.POINTER(vp)],
    "its_svo_destroy": [vp],
    "its_svo_get_info": [vp, C.POINTER(SvoInfo)],
    "its_svo_export_nodes": [vp, vp, vp, vp, vp, vp, vp, vp],
    "its_svo_export_verts": [vp, vp, vp],
    "its_svo_export_voxels": [vp, vp],
    "its_svo_export_fine_edges": [vp, vp, vp, C.c_int64, vp],
    "its_svo_export_vert_lattice": [vp, vp, vp],
    "its_sdf_at_vertices": [vp, vp, _f64p, C.c_int64, _i32p, C.c_int64, vp],
    "its_signed_distance": [vp, _f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.c_int64, _f64p],
    "its_lscg_solve": [vp, vp, vp, _f64p, C.c_double, C.c_int64, C.c_int, C.POINTER(LscgStats)],
    "its_collocation_mul": [vp, vp, _f64p, _f64p, C.c_int],
    "its_svo_set_lambda": [vp, vp, _f64p],
    "its_svo_get_lambda": [vp, _f64p],
    "its_bspline_eval": [vp, vp, _f32p, C.c_int64, _f64p, C.c_int, C.c_float, _f32p],
    "its_bspline_eval_f64pts": [vp, vp, _f64p, C.c_int64, _f64p, C.c_int, C.c_float, _f32p],
    "its_bspline_eval_device": [vp, vp, vp, vp, vp, C.c_int64, _f64p, C.c_int, C.c_float, vp, vp],
    "its_classify": [_f32p, C.c_int64, C.c_double, C.c_double, _i8p],
    "its_mc_extract": [vp, vp, vp, _f64p, _f64p, _i32p, _f64p, C.c_int, vp, C.c_int64,
                       np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.POINTER(McStats)],
    "its_svo_from_vertices": [vp, _f64p, _u32p, C.c_int64, _f64p, C.c_int64, C.POINTER(vp)],
    "its_mc_lattice": [vp, vp, _f64p, _f64p, _i32p, _f32p],
    "its_tune_set": [C.c_int, C.c_int64],
    "its_tune_get": [C.c_int, C.POINTER(C.c_int64)],
    "its_timer_start": [vp],
    "its_timer_stop": [vp, C.POINTER(C.c_double)],
    "its_lscg_diag": [vp, vp, _f64p],
    "its_lscg_structure_get": [vp, vp, C.POINTER(LscgStructure)],
    "its_shell_points": [vp, vp, _f64p, C.c_int64, _i32p, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64),
                         C.POINTER(C.c_int64)],
    "its_shell_iso": [vp, vp, _f64p, C.c_int64, _i32p, C.c_int64, _f64p, C.POINTER(C.c_int64)],
    "its_lattice_f64": [vp, vp, _f64p],
    "its_batch_create": [vp, C.POINTER(vp)],
    "its_batch_destroy": [vp],
    "its_batch_add": [vp, vp],
    "its_batch_fit": [vp, C.c_double, C.c_int64, C.c_int, vp, C.POINTER(C.c_double)],
    "its_batch_fields": [vp],
    "its_batch_eval_device": [vp, vp, vp, vp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.c_float,
                              vp, C.POINTER(C.c_double)],
    "its_batch_eval_device_noalloc": [vp, vp, vp, vp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                      C.c_float, vp, C.POINTER(C.c_double)],
    "its_mlp_create": [vp, C.c_uint64, C.POINTER(vp)],
    "its_mlp_destroy": [vp],
    "its_mlp_set_params": [vp, _f32p],
    "its_mlp_get_params": [vp, _f32p],
    "its_mlp_forward_device": [vp, vp, C.c_int64, vp, vp],
    "its_mlp_grad_device": [vp, vp, vp, C.c_int64, _f32p, C.POINTER(C.c_double)],
    "its_mlp_train_step": [vp, vp, vp, C.c_int64, C.c_float, C.POINTER(C.c_double)],
    "its_warm_features": [vp, vp, vp, vp, vp],
    "its_warm_dual": [vp, vp, _f64p, _f64p],
    "its_warm_apply": [vp, vp, _f64p, _f64p],
    "its_lscg_solve_mlp": [vp, vp, vp, vp, _f64p, C.c_double, C.c_int64, C.POINTER(LscgStats),
                           C.POINTER(C.c_double)],
    "its_hbm_bandwidth": [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)],
}
_RET = {"its_last_error": C.c_char_p, "its_version": C.c_char_p}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)"
            )
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        for name, r in _RET.items():
            fn = getattr(L, name)
            fn.argtypes = []
            fn.restype = r
        _lib = L
    return _lib


def check(rc):
    if rc != ITS_OK:
        msg = lib().its_last_error()
        raise ItsError(rc, msg.decode() if msg else "")
    return rc


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def tune_set(key, value):
    check(lib().its_tune_set(int(key), int(value)))


def tune_get(key):
    v = C.c_int64(0)
    check(lib().its_tune_get(int(key), C.byref(v)))
    return v.value


def device_count():
    n = C.c_int(0)
    check(lib().its_device_count(C.byref(n)))
    return n.value


def model_to_unit_cube(V, scale_factor=1.0):
    V = _f64(V).copy()
    check(lib().its_model_to_unit_cube(V.reshape(-1), len(V), scale_factor))
    return V


def uniform_bbox(V):
    o, e, w = np.zeros(3), np.zeros(3), np.zeros(3)
    check(lib().its_uniform_bbox(_f64(V).reshape(-1), len(V), o, e, w))
    return o, e, w


class Context:
    def __init__(self, device=0):
        h = vp()
        check(lib().its_ctx_create(device, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr):
        check(lib().its_ctx_set_stream(self._h, vp(stream_ptr)))

    def synchronize(self):
        check(lib().its_ctx_synchronize(self._h))

    def timer_start(self):
        check(lib().its_timer_start(self._h))

    def timer_stop(self):
        ms = C.c_double(0)
        check(lib().its_timer_stop(self._h, C.byref(ms)))
        return ms.value

    def hbm_bandwidth(self, nbytes=1 << 30, reps=10):
        """Achievable device bandwidth in GB/s (read + write) of a plain
        streaming copy of nbytes, median of reps launches."""
        g = C.c_double(0)
        check(lib().its_hbm_bandwidth(self._h, int(nbytes), int(reps), C.byref(g)))
        return g.value

    def close(self):
        if getattr(self, "_h", None):
            lib().its_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def signed_distance(self, V, F, pts):
        pts = _f64(pts)
        out = np.zeros(len(pts))
        check(lib().its_signed_distance(self._h, _f64(V).reshape(-1), len(V),
                                        np.ascontiguousarray(F, np.int32).reshape(-1), len(F),
                                        pts.reshape(-1), len(pts), out))
        return out


class Octree:
    """Device sparse voxel octree (SparseVoxelOctree) plus its fitted field."""

    def __init__(self, ctx: Context, tris, origin, width, R):
        tris = _f64(tris).reshape(-1, 9)
        self.ctx = ctx
        self.origin = _f64(origin)
        self.width = _f64(width)
        h = vp()
        check(lib().its_svo_build(ctx.handle, tris.reshape(-1), len(tris), self.origin, self.width, int(R),
                                  C.byref(h)))
        self._h = h
        self.info = SvoInfo()
        check(lib().its_svo_get_info(h, C.byref(self.info)))

    @classmethod
    def from_device(cls, ctx: Context, d_tris_ptr, T, origin, width, R):
        self = cls.__new__(cls)
        self.ctx = ctx
        self.origin = _f64(origin)
        self.width = _f64(width)
        h = vp()
        check(lib().its_svo_build_device(ctx.handle, vp(d_tris_ptr), int(T), self.origin, self.width, int(R),
                                         C.byref(h)))
        self._h = h
        self.info = SvoInfo()
        check(lib().its_svo_get_info(h, C.byref(self.info)))
        return self

    @classmethod
    def from_vertices(cls, ctx: Context, xyz, owner, node_width):
        """Octree object holding only a vertex set (nodeVertexArray positions
        and owners, nodeWidthArray per node as (n_nodes, 3) or (n_nodes,)):
        evaluation and marching cubes, no nodes or solver."""
        self = cls.__new__(cls)
        self.ctx = ctx
        xyz = _f64(xyz).reshape(-1, 3)
        nw = _f64(node_width)
        if nw.ndim == 1:
            nw = np.repeat(nw[:, None], 3, axis=1)
        nw = np.ascontiguousarray(nw)
        h = vp()
        check(lib().its_svo_from_vertices(ctx.handle, xyz.reshape(-1), np.ascontiguousarray(owner, np.uint32),
                                          len(xyz), nw.reshape(-1), len(nw), C.byref(h)))
        self._h = h
        self.info = SvoInfo()
        check(lib().its_svo_get_info(h, C.byref(self.info)))
        self.origin = np.array(self.info.grid_origin[:3])
        self.width = np.full(3, self.info.voxel_width * self.info.R)
        return self

    @property
    def handle(self):
        return self._h

    @property
    def depth(self):
        return self.info.tree_depth

    @property
    def num_nodes(self):
        return self.info.num_nodes

    @property
    def num_verts(self):
        return self.info.num_verts

    @property
    def num_voxels(self):
        return self.info.num_voxels

    def depth_num_nodes(self):
        return np.array(self.info.depth_num_nodes[: self.depth], np.int64)

    def esum_verts(self):
        return np.array(self.info.esum_verts[: self.depth + 1], np.int64)

    def close(self):
        if getattr(self, "_h", None):
            lib().its_svo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nodes(self):
        n = max(self.num_nodes, 1)
        out = {
            "morton": np.zeros(n, np.uint32),
            "is_leaf": np.zeros(n, np.uint8),
            "origin": np.zeros((n, 3)),
            "width": np.zeros(n),
            "parent": np.zeros(n, np.uint32),
            "childs": np.zeros((n, 8), np.uint32),
            "neighbors": np.zeros((n, 27), np.uint32),
        }
        ptr = lambda a: a.ctypes.data_as(vp)
        check(lib().its_svo_export_nodes(self._h, *(ptr(out[k]) for k in
                                                    ["morton", "is_leaf", "origin", "width", "parent", "childs",
                                                     "neighbors"])))
        return {k: v[: self.num_nodes] for k, v in out.items()}

    def verts(self):
        n = max(self.num_verts, 1)
        xyz = np.zeros((n, 3))
        owner = np.zeros(n, np.uint32)
        check(lib().its_svo_export_verts(self._h, xyz.ctypes.data_as(vp), owner.ctypes.data_as(vp)))
        return xyz[: self.num_verts], owner[: self.num_verts]

    def voxels(self):
        out = np.zeros(max(self.num_voxels, 1), np.uint32)
        check(lib().its_svo_export_voxels(self._h, out.ctypes.data_as(vp)))
        return out[: self.num_voxels]

    def vert_lattice(self):
        """Per vertex: lattice coordinates at its level (n, 3) and the level."""
        n = max(self.num_verts, 1)
        ijk = np.zeros((n, 3), np.uint16)
        lev = np.zeros(n, np.uint8)
        check(lib().its_svo_export_vert_lattice(self._h, ijk.ctypes.data_as(vp), lev.ctypes.data_as(vp)))
        return ijk[: self.num_verts], lev[: self.num_verts]

    def fine_edges(self):
        """Level-0 node edges: (n, 2, 3) end points and owner node per edge."""
        n = C.c_int64(0)
        check(lib().its_svo_export_fine_edges(self._h, None, None, 0, C.byref(n)))
        ends = np.zeros((max(n.value, 1), 2, 3))
        owner = np.zeros(max(n.value, 1), np.uint32)
        check(lib().its_svo_export_fine_edges(self._h, ends.ctypes.data_as(vp), owner.ctypes.data_as(vp), n.value,
                                              C.byref(n)))
        return ends[: n.value], owner[: n.value]

    def sdf_at_vertices(self, V, F):
        b = np.zeros(max(self.num_verts, 1))
        check(lib().its_sdf_at_vertices(self.ctx.handle, self._h, _f64(V).reshape(-1), len(V),
                                        np.ascontiguousarray(F, np.int32).reshape(-1), len(F),
                                        b.ctypes.data_as(vp)))
        return b[: self.num_verts]

    def lscg(self, b=None, x0=None, tol=1e-6, max_iter=0):
        lam = np.zeros(self.num_verts) if x0 is None else _f64(x0).copy()
        st = LscgStats()
        bp = _f64(b).ctypes.data_as(vp) if b is not None else None
        bb = _f64(b) if b is not None else None
        check(lib().its_lscg_solve(self.ctx.handle, self._h, bb.ctypes.data_as(vp) if bb is not None else None, lam,
                                   tol, int(max_iter), 0 if x0 is None else 1, C.byref(st)))
        del bp
        return lam, st.as_dict()

    def lscg_mlp(self, net, b=None, tol=1e-6, max_iter=0):
        """Warm-started solve: features -> network -> lambda0 = D A^T w -> LSCG."""
        lam = np.zeros(self.num_verts)
        st = LscgStats()
        tp = C.c_double(0)
        bb = _f64(b) if b is not None else None
        check(lib().its_lscg_solve_mlp(self.ctx.handle, self._h, net.handle,
                                       bb.ctypes.data_as(vp) if bb is not None else None, lam, tol, int(max_iter),
                                       C.byref(st), C.byref(tp)))
        d = st.as_dict()
        d["seconds_predict"] = tp.value
        return lam, d

    def shell_points(self, V, F):
        """(points, n_edge): mesh-edge x node-face points, then node-edge x face points."""
        V, F = _f64(V), np.ascontiguousarray(F, np.int32)
        ne, nf = C.c_int64(0), C.c_int64(0)
        check(lib().its_shell_points(self.ctx.handle, self._h, V.reshape(-1), len(V), F.reshape(-1), len(F), None, 0,
                                     C.byref(ne), C.byref(nf)))
        n = ne.value + nf.value
        out = np.zeros((max(n, 1), 3))
        check(lib().its_shell_points(self.ctx.handle, self._h, V.reshape(-1), len(V), F.reshape(-1), len(F),
                                     out.ctypes.data_as(vp), n, C.byref(ne), C.byref(nf)))
        return out[:n], ne.value

    def shell_iso(self, V, F):
        """(inner, outer) iso values and the number of intersection points."""
        V, F = _f64(V), np.ascontiguousarray(F, np.int32)
        iso = np.zeros(2)
        n = C.c_int64(0)
        check(lib().its_shell_iso(self.ctx.handle, self._h, V.reshape(-1), len(V), F.reshape(-1), len(F), iso,
                                  C.byref(n)))
        return float(iso[0]), float(iso[1]), n.value

    def lscg_structure(self):
        """n, assembled non-zeros, parent entries and bytes per CGLS iteration."""
        out = LscgStructure()
        check(lib().its_lscg_structure_get(self.ctx.handle, self._h, C.byref(out)))
        return out.as_dict()

    def lscg_diag(self):
        d = np.zeros(self.num_verts)
        check(lib().its_lscg_diag(self.ctx.handle, self._h, d))
        return d

    def lattice_f64(self):
        D = self.info.R + 1
        out = np.zeros(D * D * D)
        check(lib().its_lattice_f64(self.ctx.handle, self._h, out))
        return out

    def warm_features(self, d_x, d_scale, stream=None):
        check(lib().its_warm_features(self.ctx.handle, self._h, vp(d_x), vp(d_scale), vp(stream) if stream else None))

    def warm_dual(self, lam):
        w = np.zeros(self.num_verts)
        check(lib().its_warm_dual(self.ctx.handle, self._h, _f64(lam), w))
        return w

    def warm_apply(self, w):
        lam0 = np.zeros(self.num_verts)
        check(lib().its_warm_apply(self.ctx.handle, self._h, _f64(w), lam0))
        return lam0

    def mul(self, x, transpose=False):
        y = np.zeros(self.num_verts)
        check(lib().its_collocation_mul(self.ctx.handle, self._h, _f64(x), y, 1 if transpose else 0))
        return y

    def set_lambda(self, lam):
        check(lib().its_svo_set_lambda(self.ctx.handle, self._h, _f64(lam)))

    def get_lambda(self):
        out = np.zeros(self.num_verts)
        check(lib().its_svo_get_lambda(self._h, out))
        return out

    def eval(self, pts, box_end, mode=EVAL_SUPPORT, out_val=0.0):
        pts = np.asarray(pts)
        f = np.zeros(len(pts), np.float32)
        if pts.dtype == np.float64:
            check(lib().its_bspline_eval_f64pts(self.ctx.handle, self._h, np.ascontiguousarray(pts).reshape(-1),
                                                len(pts), _f64(box_end), mode, out_val, f))
        else:
            check(lib().its_bspline_eval(self.ctx.handle, self._h,
                                         np.ascontiguousarray(pts, np.float32).reshape(-1), len(pts), _f64(box_end),
                                         mode, out_val, f))
        return f

    def eval_device(self, dx, dy, dz, N, box_end, d_out, mode=EVAL_SUPPORT, out_val=0.0, stream=None):
        check(lib().its_bspline_eval_device(self.ctx.handle, self._h, vp(dx), vp(dy), vp(dz), int(N), _f64(box_end),
                                            mode, out_val, vp(d_out), vp(stream) if stream else None))

    def marching_cubes(self, origin, width, res, iso, lam=None):
        """Soup (num_verts, 3) binary32 and stats.  The count call evaluates
        the lattice; the fill call reuses it, so the lattice times come from
        the first call and the extraction times from the second.  lam: new
        coefficients to extract from (None: the current ones)."""
        soups, st = self.marching_cubes_multi(origin, width, res, [iso], lam)
        return soups[0], st

    def marching_cubes_multi(self, origin, width, res, isos, lam=None):
        """One soup per iso value from one lattice: ([soup_k], stats)."""
        res = np.asarray(res, np.int32)
        isos = _f64(np.atleast_1d(isos))
        counts = np.zeros(len(isos), np.int64)
        st = McStats()
        lam_keep = _f64(lam) if lam is not None else None
        lp = lam_keep.ctypes.data_as(vp) if lam_keep is not None else None
        check(lib().its_mc_extract(self.ctx.handle, self._h, lp, _f64(origin), _f64(width), res, isos, len(isos),
                                   None, 0, counts, C.byref(st)))
        n = int(counts.sum())
        lat = (st.seconds_eval, st.seconds_eval_device)
        tri = np.zeros((max(n, 1), 3), np.float32)
        if n:
            check(lib().its_mc_extract(self.ctx.handle, self._h, None, _f64(origin), _f64(width), res, isos,
                                       len(isos), tri.ctypes.data_as(vp), n, counts, C.byref(st)))
        offs = np.concatenate([[0], np.cumsum(counts)])
        soups = [tri[offs[k]:offs[k + 1]] for k in range(len(isos))]
        return soups, {"num_verts": st.num_verts, "num_valid_voxels": st.num_valid_voxels,
                       "seconds_eval": lat[0], "seconds_extract": st.seconds_extract,
                       "seconds_eval_device": lat[1],
                       "seconds_extract_device": st.seconds_extract_device, "counts": counts}

    def mc_lattice(self, origin, width, res):
        res = np.asarray(res, np.int32)
        out = np.zeros(int(np.prod(res + 1)), np.float32)
        check(lib().its_mc_lattice(self.ctx.handle, self._h, _f64(origin), _f64(width), res, out))
        return out


class Batch:
    """Octrees fitted and evaluated together (one launch each)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = vp()
        check(lib().its_batch_create(ctx.handle, C.byref(h)))
        self._h = h
        self.octrees = []

    def add(self, svo):
        check(lib().its_batch_add(self._h, svo.handle))
        self.octrees.append(svo)

    def fit(self, tol=1e-6, max_iter=0, warm=False):
        n = len(self.octrees)
        st = (LscgStats * n)()
        sec = C.c_double(0)
        check(lib().its_batch_fit(self._h, tol, int(max_iter), 1 if warm else 0, C.cast(st, vp), C.byref(sec)))
        return [s.as_dict() for s in st], sec.value

    def fields(self):
        check(lib().its_batch_fields(self._h))

    def eval_device(self, dx, dy, dz, offsets, d_out, out_val=0.0):
        sec = C.c_double(0)
        check(lib().its_batch_eval_device(self._h, vp(dx), vp(dy), vp(dz), np.ascontiguousarray(offsets, np.int64),
                                          float(out_val), vp(d_out), C.byref(sec)))
        return sec.value

    def eval_device_noalloc(self, dx, dy, dz, offsets, d_out, out_val=0.0):
        """eval_device with the library's fixed scratch array: no allocation."""
        sec = C.c_double(0)
        check(lib().its_batch_eval_device_noalloc(self._h, vp(dx), vp(dy), vp(dz),
                                                  np.ascontiguousarray(offsets, np.int64), float(out_val), vp(d_out),
                                                  C.byref(sec)))
        return sec.value

    def close(self):
        if getattr(self, "_h", None):
            lib().its_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_IN, MLP_HIDDEN, MLP_NPARAM = 16, 64, 9473


class Mlp:
    """Warm-start network (BF16 MFMA kernels); parameters as binary32."""

    def __init__(self, ctx: Context, seed=0):
        self.ctx = ctx
        h = vp()
        check(lib().its_mlp_create(ctx.handle, int(seed), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().its_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        check(lib().its_mlp_get_params(self._h, p))
        return p

    def set_params(self, p):
        check(lib().its_mlp_set_params(self._h, np.ascontiguousarray(p, np.float32)))

    def forward_device(self, d_x, rows, d_y, stream=None):
        check(lib().its_mlp_forward_device(self._h, vp(d_x), int(rows), vp(d_y), vp(stream) if stream else None))

    def grad_device(self, d_x, d_t, rows):
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double(0)
        check(lib().its_mlp_grad_device(self._h, vp(d_x), vp(d_t), int(rows), g, C.byref(loss)))
        return g, loss.value

    def train_step(self, d_x, d_t, rows, lr):
        loss = C.c_double(0)
        check(lib().its_mlp_train_step(self._h, vp(d_x), vp(d_t), int(rows), float(lr), C.byref(loss)))
        return loss.value


def classify(f, inner, outer):
    f = np.ascontiguousarray(f, np.float32)
    out = np.zeros(len(f), np.int8)
    check(lib().its_classify(f, len(f), inner, outer, out))
    return out
