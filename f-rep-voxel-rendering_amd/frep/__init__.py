"""MI355X F-rep voxelizer / raytracer: Python side of the C ABI (include/frep.h).

``frep.tape`` / ``frep.shapes`` build tapes in the reference's format;
``frep.lib`` binds libfrep.so (HIP kernels for gfx950) through ctypes.
"""
from . import shapes, tape  # noqa: F401
