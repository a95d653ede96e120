"""its_amd: MI355X (gfx950) implementation of the ITS hot path.

The sparse voxel octree build, hat B-spline evaluation, LSCG coefficient fit
and marching cubes are hand-written HIP kernels.  They live in libits_hip.so
behind the C ABI of include/its_c.h.  This package holds the ctypes binding
(:mod:`its_amd._capi`), the seeded benchmark meshes (:mod:`its_amd.meshes`),
the batching, autotuning and warm-start drivers, and, under its_amd/host/,
the C++ facade with the reference's ThinShells, SparseVoxelOctree and
MC::marching_cubes signatures.
"""
from . import _capi  # noqa: F401
from ._capi import Context, Octree, LibraryMissing, ItsError  # noqa: F401

__all__ = ["Context", "Octree", "LibraryMissing", "ItsError"]
