"""MI355X-native ray traversal for the Vermeer renderer's hot path.

The work is done by libvermeer_gpu.so (HIP kernels for gfx950 + a C++ host
builder) behind the C ABI declared in include/vermeer_gpu.h; this package is
the Python mirror of the reference's plugin interface over that ABI.
"""
from . import _lib
from .accel import PolyMesh, Scene, build_accel, build_accel_gpu, build_accel_motion
from ._lib import HIT_DTYPE, MNODE_DTYPE, NODE_DTYPE, RAY_DTYPE, VgpuError

__all__ = ["PolyMesh", "Scene", "build_accel", "build_accel_gpu", "build_accel_motion", "HIT_DTYPE", "NODE_DTYPE", "MNODE_DTYPE",
           "RAY_DTYPE", "VgpuError"]
