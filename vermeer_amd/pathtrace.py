"""Path mode: full paths with next-event estimation and optional guiding (no reference counterpart).

Thin mirror of the vgpu_pt_* entry points of include/vermeer_gpu.h.  The scene description is the
one vermeer_amd.scenes.guided_room_scene() returns; materials are vgpu_pt_material records.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import CAMERA_DTYPE, check, ptr
from .accel import Scene
from .render import lookat_camera
from .scenes import PT_MATERIAL_DTYPE


def camera_for(desc, width, height):
    c = desc["camera"]
    return lookat_camera(c["from"], c["to"], c["up"], c["fov"], c["focal"], width, height)


def build_scene(desc, device=0, leaf_size=16):
    """Committed GPU scene of the description's shaded meshes (shareable by several PathTracers).
    leaf_size: BVH leafMax of the meshes (16, the reference's, keeps the reference's hit IDs)."""
    lib = _lib.load()
    scene = Scene(device)
    scene.set_leaf_size(leaf_size)
    for mesh in desc["meshes"]:
        v = np.ascontiguousarray(mesh["verts"], np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(mesh["idx"], np.uint32).reshape(-1, 3)
        sidx = mesh.get("shaderidx")
        sidx = None if sidx is None else np.ascontiguousarray(sidx, np.uint8)
        sh = np.ascontiguousarray(mesh["shaders"], np.int32)
        g = C.c_int32(-1)
        check(lib.vgpu_scene_add_shaded_mesh(scene._h, ptr(v), len(v), ptr(ix), len(ix), None, ptr(sidx), ptr(sh),
                                             len(sh), C.byref(g)))
        scene.geoms.append(None)
    scene.pre_render()
    return scene


def scene_bounds(desc, margin=0.05):
    lo = np.min([np.asarray(m["verts"]).reshape(-1, 3).min(0) for m in desc["meshes"]], axis=0) - margin
    hi = np.max([np.asarray(m["verts"]).reshape(-1, 3).max(0) for m in desc["meshes"]], axis=0) + margin
    return np.concatenate([lo, hi]).astype(np.float32)


class PathTracer:
    def __init__(self, desc, width, height, max_depth=8, seed=1, cam=None, device=0, scene=None, leaf_size=16):
        self._lib = _lib.load()
        self._own_scene = scene is None
        self.scene = build_scene(desc, device, leaf_size) if scene is None else scene
        self.width, self.height = width, height
        mats = np.ascontiguousarray(desc["materials"], PT_MATERIAL_DTYPE)
        self.cam = np.ascontiguousarray(cam if cam is not None else camera_for(desc, width, height), CAMERA_DTYPE)
        h = C.c_void_p()
        check(self._lib.vgpu_pt_create(C.byref(h), self.scene._h, ptr(mats), len(mats), ptr(self.cam), width, height,
                                       max_depth, int(seed)))
        self._h = h
        self.guide = None
        self.accum = np.zeros((height, width, 3), np.float64)
        self.passes = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vgpu_pt_destroy(self._h)
            self._h = None
        if self._own_scene:
            self.scene.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_guide(self, guide, alpha=0.5, record_stride=16, train_steps=4, batch=65536, lr=1e-3):
        self.guide = guide
        gh = guide._h if guide is not None else None
        check(self._lib.vgpu_pt_set_guide(self._h, gh, float(alpha), int(record_stride), int(train_steps), int(batch),
                                          float(lr)))

    def render_pass(self, pass_index, train=True):
        flags = 0 if train else _lib.VGPU_PT_NO_TRAIN
        check(self._lib.vgpu_pt_pass(self._h, flags, int(pass_index), ptr(self.accum), None))
        self.passes += 1

    def render_pass_device(self, d_accum, pass_index, stream=None, train=True, graph=False):
        """One pass on a device accumulation buffer; graph=True replays the pass as a HIP graph."""
        flags = (_lib.VGPU_TRACE_DEVICE_PTRS | (0 if train else _lib.VGPU_PT_NO_TRAIN) |
                 (_lib.VGPU_GRAPH if graph else 0))
        check(self._lib.vgpu_pt_pass(self._h, flags, int(pass_index), C.c_void_p(d_accum), C.c_void_p(stream)))

    def stats(self):
        rays = C.c_uint64()
        ts = C.c_uint32()
        check(self._lib.vgpu_pt_stats(self._h, C.byref(rays), C.byref(ts)))
        return int(rays.value), int(ts.value)

    @property
    def image(self):
        return self.accum / max(self.passes, 1)
