"""ctypes binding of libvermeer_gpu.so (the C ABI in include/vermeer_gpu.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvermeer_gpu.so")

VGPU_OK = 0
VGPU_ERR_ARG = -1
VGPU_ERR_HIP = -2
VGPU_ERR_BUILD = -3
VGPU_ERR_STACK = -4
VGPU_ERR_STATE = -5
VGPU_ERR_NOMEM = -6
VGPU_MAX_STACK = 128

VGPU_TRACE_SHADOW = 1
VGPU_TRACE_DEVICE_PTRS = 2
VGPU_MESH_MOTION_REFERENCE_INDEXING = 1
VGPU_GUIDE_RAW = 4
VGPU_GUIDE_NO_UPDATE = 8
VGPU_GUIDE_NPARAMS = 24360
VGPU_PT_NO_TRAIN = 16
VGPU_GRAPH = 32
VGPU_LIGHT_TRI = 0
VGPU_LIGHT_DISK = 1
VGPU_LIGHT_SPHERE = 2
VGPU_FILTER_BOX = 0
VGPU_FILTER_GAUSSIAN = 1
VGPU_FILTER_AIRY = 2
VGPU_VNF_MESSAGES = 0
VGPU_VNF_NODES = 1
VGPU_OUTPUT_HDR = 0
VGPU_OUTPUT_FLOAT = 1
VGPU_DIAG_COPY = 0
VGPU_DIAG_READ = 1

# vgpu_qbvh_node (128 bytes, qbvh.Node layout)
NODE_DTYPE = np.dtype([("boxes", "<f4", 24), ("axis", "<u4", 3), ("children", "<i4", 4), ("parent", "<i4")])
MNODE_DTYPE = np.dtype([("axis", "<i4", 3), ("children", "<i4", 4), ("pad", "<i4")])
RAY_DTYPE = np.dtype([("P", "<f4", 3), ("tmax", "<f4"), ("D", "<f4", 3), ("time", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
assert NODE_DTYPE.itemsize == 128 and RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and MNODE_DTYPE.itemsize == 32


class TraceConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("lds_stack", C.c_int32), ("waves_per_cu", C.c_int32), ("sort", C.c_int32),
                ("refill", C.c_int32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# (name, restype, argtypes) for every entry point declared in include/vermeer_gpu.h
_P = C.c_void_p
_I32 = C.c_int32
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float
SIGNATURES = [
    ("vgpu_last_error", C.c_char_p, []),
    ("vgpu_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("vgpu_qbvh_build", C.c_int, [_P, _P, _P, _I32, _I32, _P, _I32, C.POINTER(_I32), _P, _I32]),
    ("vgpu_qbvh_build_gpu", C.c_int, [_P, _P, _P, _I32, _I32, _P, _I32, C.POINTER(_I32), _P, _P]),
    ("vgpu_qbvh_build_motion", C.c_int, [_P, _P, _P, _I32, _I32, _P, _I32, C.POINTER(_I32), _I32]),
    ("vgpu_scene_create", C.c_int, [C.POINTER(_P), C.c_int]),
    ("vgpu_scene_destroy", C.c_int, [_P]),
    ("vgpu_scene_add_mesh", C.c_int, [_P, _P, _U32, _U32, _P, _U32, _F, _U32, C.POINTER(_I32)]),
    ("vgpu_scene_add_mesh_prebuilt", C.c_int, [_P, _P, _U32, _P, _U32, _P, _U32, _P, _F, C.POINTER(_I32)]),
    ("vgpu_scene_add_motion_mesh_prebuilt", C.c_int,
     [_P, _P, _U32, _U32, _P, _U32, _P, _P, _U32, _P, _F, _U32, C.POINTER(_I32)]),
    ("vgpu_scene_commit", C.c_int, [_P]),
    ("vgpu_scene_info", C.c_int, [_P, C.POINTER(_I32), C.POINTER(_I32), _P, _P, _P]),
    ("vgpu_scene_mesh_accel", C.c_int, [_P, _I32, _P, _P, _P]),
    ("vgpu_trace", C.c_int, [_P, _U32, _P, _U64, _P, _P, _P, _P]),
    ("vgpu_scene_status", C.c_int, [_P, _P]),
    ("vgpu_trace_noalloc", C.c_int, [_P, _U32, _P, _U64, _P, _P, _P, _P, _P]),
    ("vgpu_scene_set_trace_config", C.c_int, [_P, C.POINTER(TraceConfig)]),
    ("vgpu_scene_get_trace_config", C.c_int, [_P, C.POINTER(TraceConfig)]),
    ("vgpu_scene_autotune", C.c_int, [_P, _U32, _P, _U64, _P, _P, C.POINTER(TraceConfig), C.POINTER(_F)]),
    ("vgpu_scene_set_leaf_size", C.c_int, [_P, _I32]),
    ("vgpu_scene_set_timing", C.c_int, [_P, _I32]),
    ("vgpu_scene_timing", C.c_int, [_P, C.POINTER(_F), C.POINTER(_F), C.POINTER(_I32)]),
    ("vgpu_diag_stream", C.c_int, [_I32, _P, _P, _U64, _I32, _P]),
    ("vgpu_film_to_rgbe", C.c_int, [C.c_int, _U32, _P, _U32, _U32, _P, _P]),
    ("vgpu_vnf_load", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("vgpu_vnf_destroy", C.c_int, [_P]),
    ("vgpu_vnf_info", C.c_int, [_P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_I32), C.POINTER(_I32)]),
    ("vgpu_vnf_text", C.c_int, [_P, _I32, C.c_char_p, _U64, C.POINTER(_U64)]),
    ("vgpu_vnf_build", C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(_P)]),
    ("vgpu_vnf_output", C.c_int, [_P, _I32, C.POINTER(_I32), C.c_char_p, _U64, C.POINTER(_I32)]),
    ("vgpu_write_hdr", C.c_int, [C.c_int, _U32, _P, _U32, _U32, C.c_char_p, _P]),
    ("vgpu_write_float", C.c_int, [C.c_int, _U32, _P, _U32, _U32, C.c_char_p, _P]),
    ("vgpu_scene_add_shaded_mesh", C.c_int, [_P, _P, _U32, _P, _U32, _P, _P, _P, _U32, C.POINTER(_I32)]),
    ("vgpu_scene_add_shaded_motion_mesh", C.c_int,
     [_P, _P, _U32, _U32, _P, _U32, _P, _U32, _P, _P, _U32, _U32, C.POINTER(_I32)]),
    ("vgpu_scene_add_sphere", C.c_int, [_P, _P, _F, _I32, C.POINTER(_I32)]),
    ("vgpu_render_create", C.c_int, [C.POINTER(_P), _P, _P, _U32, _P, _U32, _P, _U32, _U32]),
    ("vgpu_render_set_filter", C.c_int, [_P, _I32, _F, _I32, _F]),
    ("vgpu_render_set_path_depth", C.c_int, [_P, C.c_uint32]),
    ("vgpu_scene_set_transform", C.c_int, [_P, _I32, _P, C.c_uint32]),
    ("vgpu_scene_set_hidden", C.c_int, [_P, _I32, _I32]),
    ("vgpu_scene_add_instance", C.c_int, [_P, _I32, _P, C.c_uint32, _P, _P, _P]),
    ("vgpu_scene_add_proc", C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _P, _P, _P]),
    ("vgpu_render_stats", C.c_int, [_P, _P, _P]),
    ("vgpu_render_set_camera_keys", C.c_int, [_P, _P, _U32]),
    ("vgpu_render_destroy", C.c_int, [_P]),
    ("vgpu_render_pass", C.c_int, [_P, _U32, _P, _U32, _P, _P, _P, C.POINTER(_U64)]),
    ("vgpu_guide_create", C.c_int, [C.POINTER(_P), C.c_int, _P, _U64, _U32]),
    ("vgpu_guide_destroy", C.c_int, [_P]),
    ("vgpu_guide_set_params", C.c_int, [_P, _P]),
    ("vgpu_guide_get_params", C.c_int, [_P, _P]),
    ("vgpu_guide_get_grad", C.c_int, [_P, _P]),
    ("vgpu_guide_infer", C.c_int, [_P, _U32, _P, _U64, _P, _P]),
    ("vgpu_guide_train_step", C.c_int, [_P, _U32, _P, _U32, _F, _P, C.POINTER(_F)]),
    ("vgpu_pt_create", C.c_int, [C.POINTER(_P), _P, _P, _U32, _P, _U32, _U32, _U32, _U64]),
    ("vgpu_pt_destroy", C.c_int, [_P]),
    ("vgpu_pt_set_guide", C.c_int, [_P, _P, _F, _U32, _U32, _U32, _F]),
    ("vgpu_pt_pass", C.c_int, [_P, _U32, _U32, _P, _P]),
    ("vgpu_pt_stats", C.c_int, [_P, C.POINTER(_U64), C.POINTER(_U32)]),
]

# vgpu_material / vgpu_light / vgpu_camera layouts
MATERIAL_DTYPE = np.dtype([
    ("emission", "<f4", 3), ("emission_strength", "<f4"), ("has_emission_colour", "<i4"), ("has_emission_strength", "<i4"),
    ("diffuse", "<f4", 3), ("has_diffuse_colour", "<i4"),
    ("diffuse_strength", "<f4"), ("has_diffuse_strength", "<i4"),
    ("diffuse_roughness", "<f4"), ("has_diffuse_roughness", "<i4"),
    ("spec1", "<f4", 3), ("has_spec1_colour", "<i4"),
    ("spec1_strength", "<f4"), ("has_spec1_strength", "<i4"),
    ("spec1_roughness", "<f4"), ("has_spec1_roughness", "<i4"),
    ("ior", "<f4"), ("has_ior", "<i4"),
    ("fresnel_model", "<i4"),
    ("spec1_fresnel_refl", "<f4", 3), ("has_spec1_fresnel_refl", "<i4"),
    ("spec1_fresnel_edge", "<f4", 3), ("has_spec1_fresnel_edge", "<i4"),
])
LIGHT_DTYPE = np.dtype([("P0", "<f4", 3), ("P1", "<f4", 3), ("P2", "<f4", 3), ("material", "<i4"),
                        ("samples", "<i4"), ("geom", "<i4"), ("type", "<i4"), ("radius", "<f4")])
CAMERA_DTYPE = np.dtype([("M", "<f4", 16), ("tan_theta_focal", "<f4"), ("aspect", "<f4"), ("focal", "<f4"),
                         ("radius", "<f4")])

_lib = None


class VgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vgpu error {code}: {msg}")
        self.code = code


def load():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                           " (there is no CPU fallback)")
    # One HIP runtime per process.  torch ships its own libamdhip64 (SONAME libamdhip64.so.7) and
    # loads it by a different file name; loading this library first would pull in /opt/rocm's copy
    # and leave two runtimes fighting over the device.  Importing torch first makes the dynamic
    # linker bind this library to the runtime torch already holds.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != VGPU_OK:
        raise VgpuError(rc, load().vgpu_last_error().decode())
    return rc


def ptr(a):
    """Raw data pointer of a numpy array (or None)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)
