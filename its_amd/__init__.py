"""its_amd: MI355X (gfx950) implementation of the ITS hot path.

The sparse voxel octree build, hat B-spline evaluation, LSCG coefficient fit
and marching cubes are hand-written HIP kernels.  They live in libits_hip.so
behind the C ABI of include/its_c.h.  This package holds the ctypes binding
(:mod:`its_amd._capi`), the seeded benchmark meshes (:mod:`its_amd.meshes`)
and the ThinShells-style facade (:mod:`its_amd.thinshells`).
"""
from . import _capi  # noqa: F401
from ._capi import Context, Octree, LibraryMissing, ItsError  # noqa: F401

__all__ = ["Context", "Octree", "LibraryMissing", "ItsError"]
