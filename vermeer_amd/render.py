"""Renderer: the reference's render pass loop on the GPU.

Mirrors core.Render (core/render.go:140-218) for a scene of shaded meshes (static or
moving), analytic spheres and Tri / Disk / Sphere lights: create once per frame
setup, then one pass() per iteration with the 1-based pass number the
reference hands to render().
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import CAMERA_DTYPE, LIGHT_DTYPE, MATERIAL_DTYPE, check, ptr
from .accel import Scene


def lookat_camera(frm, to, up, fov_deg, focal, width, height):
    """Camera record for a LookAt camera (camera.go:109-193), computed directly from the basis.

    The reference recovers the same matrix through a polar decomposition and quaternion round trip,
    which can differ in the last bit; callers that need the reference's exact matrix pass it in."""
    f = np.asarray(frm, np.float32)
    t = np.asarray(to, np.float32)
    u = np.asarray(up, np.float32)
    W = (f - t) / np.linalg.norm(f - t)
    U = np.cross(u, W)
    U = U / np.linalg.norm(U)
    V = np.cross(U, W)
    V = V / np.linalg.norm(V)
    cam = np.zeros((), CAMERA_DTYPE)
    M = np.eye(4, dtype=np.float32)
    M[0, :3], M[1, :3], M[2, :3], M[3, :3] = U, V, W, f      # column-major storage: columns U, V, W, T
    cam["M"] = M.reshape(16)
    cam["tan_theta_focal"] = np.float32(math.tan(math.radians(fov_deg / 2.0)) * focal)
    cam["aspect"] = np.float32(width / height)
    cam["focal"] = np.float32(focal)
    return cam


_LIGHT_TYPES = {"tri": _lib.VGPU_LIGHT_TRI, "disk": _lib.VGPU_LIGHT_DISK, "sphere": _lib.VGPU_LIGHT_SPHERE}


def light_records(lights):
    """vgpu_light records from scene-description light dicts: type "tri" (P0, P1, P2), "disk" (P, Up,
    LookAt, radius) or "sphere" (P, radius); material, samples, geom (insertion index)."""
    out = np.zeros(len(lights), LIGHT_DTYPE)
    for i, L in enumerate(lights):
        kind = L.get("type", "tri")
        out[i]["type"] = _LIGHT_TYPES[kind]
        if kind == "tri":
            out[i]["P0"], out[i]["P1"], out[i]["P2"] = L["P0"], L["P1"], L["P2"]
        elif kind == "disk":
            out[i]["P0"], out[i]["P1"], out[i]["P2"] = L["P"], L["Up"], L["LookAt"]
        else:
            out[i]["P0"] = L["P"]
        if kind != "tri":
            out[i]["radius"] = L["radius"]
        out[i]["material"], out[i]["samples"], out[i]["geom"] = L["material"], L["samples"], L["geom"]
    return out


class Renderer:
    """GPU render passes over a scene description (see vermeer_amd.scenes.box_spheres_scene)."""

    def __init__(self, desc, width, height, camera, device=0):
        self._lib = _lib.load()
        self.scene = Scene(device)
        self.width, self.height = width, height
        self._keep = []
        for mesh in desc["meshes"]:
            if mesh.get("instance") is not None:   # an instance or proc in scene order
                it = mesh["instance"]
                self.scene.add_instance(it["geom"], it["matrices"], it["bmin"], it["bmax"])
                continue
            if mesh.get("proc") is not None:
                pr = mesh["proc"]
                self.scene.add_proc(pr["geoms"], pr.get("matrices"), pr["bmin"], pr["bmax"])
                continue
            if mesh.get("sphere"):
                self.scene.add_sphere(mesh["centre"], mesh["radius"], mesh["material"])
                continue
            if int(mesh.get("keys", 1)) > 1:
                self._add_motion_mesh(mesh)
                continue
            v = np.ascontiguousarray(mesh["verts"], np.float32).reshape(-1, 3)
            ix = np.ascontiguousarray(mesh["idx"], np.uint32).reshape(-1, 3)
            nrm = mesh.get("normals")
            nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
            sidx = mesh.get("shaderidx")
            sidx = None if sidx is None else np.ascontiguousarray(sidx, np.uint8)
            sh = np.ascontiguousarray(mesh["shaders"], np.int32)
            g = C.c_int32(-1)
            check(self._lib.vgpu_scene_add_shaded_mesh(self.scene._h, ptr(v), len(v), ptr(ix), len(ix), ptr(nrm),
                                                       ptr(sidx), ptr(sh), len(sh), C.byref(g)))
            self.scene.geoms.append(None)
        self.scene.add_transforms(desc)
        self.scene.pre_render()
        mats = np.ascontiguousarray(desc["materials"], MATERIAL_DTYPE)
        lights = light_records(desc["lights"])
        cam = np.ascontiguousarray(camera, CAMERA_DTYPE)
        h = C.c_void_p()
        check(self._lib.vgpu_render_create(C.byref(h), self.scene._h, ptr(mats), len(mats), ptr(lights), len(lights),
                                           ptr(cam), width, height))
        self._h = h
        self.film = np.zeros((height, width, 3), np.float32)

    @classmethod
    def adopt(cls, scene, handle, width, height):
        """A Renderer owning a vgpu_render the library created (vgpu_vnf_build) over `scene`."""
        r = cls.__new__(cls)
        r._lib = _lib.load()
        r.scene = scene
        r.width, r.height = width, height
        r._keep = []
        r._h = handle
        r.film = np.zeros((height, width, 3), np.float32)
        return r

    def _add_motion_mesh(self, mesh):
        """A moving shaded mesh: verts (keys, nverts, 3); normals (normal keys, ntris * 3, 3) per face
        corner or None; shaderidx, shaders; reference_indexing selects the reference's slot indexing."""
        nk = int(mesh["keys"])
        v = np.ascontiguousarray(mesh["verts"], np.float32).reshape(nk, -1, 3)
        ix = np.ascontiguousarray(mesh["idx"], np.uint32).reshape(-1, 3)
        nrm = mesh.get("normals")
        nrm_keys = 0
        if nrm is not None:
            nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, len(ix) * 3, 3)
            nrm_keys = nrm.shape[0]
        sidx = mesh.get("shaderidx")
        sidx = None if sidx is None else np.ascontiguousarray(sidx, np.uint8)
        sh = np.ascontiguousarray(mesh["shaders"], np.int32)
        flags = _lib.VGPU_MESH_MOTION_REFERENCE_INDEXING if mesh.get("reference_indexing") else 0
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_shaded_motion_mesh(self.scene._h, ptr(v), v.shape[1], nk, ptr(ix), len(ix),
                                                          ptr(nrm), nrm_keys, ptr(sidx), ptr(sh), len(sh), flags,
                                                          C.byref(g)))
        self.scene.geoms.append(None)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vgpu_render_destroy(self._h)
            self._h = None
        self.scene.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_path_depth(self, depth):
        """Path mode: paths of up to `depth` segments shaded by ShaderStd at every vertex, one
        continuation ray per vertex (vgpu_render_set_path_depth); 0 keeps the reference's shading."""
        check(self._lib.vgpu_render_set_path_depth(self._h, int(depth)))

    def stats(self):
        """(closest-hit rays, shadow rays) traced since creation (vgpu_render_stats; synchronises)."""
        a, b = C.c_uint64(0), C.c_uint64(0)
        check(self._lib.vgpu_render_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_filter(self, kind="box", width=0.0, res=0, peak=0.0):
        """Pixel filter: "box" (none), "gaussian" or "airy"; zeros take the node defaults."""
        t = {"box": _lib.VGPU_FILTER_BOX, "gaussian": _lib.VGPU_FILTER_GAUSSIAN, "airy": _lib.VGPU_FILTER_AIRY}[kind]
        check(self._lib.vgpu_render_set_filter(self._h, t, float(width), int(res), float(peak)))

    def set_camera_keys(self, local_to_world):
        """Camera motion keys: (nkeys, 16) column-major LocalToWorld matrices; each sample's camera matrix
        is recomposed from the lerp of the two keys bracketing its time (one key: a single matrix)."""
        m = np.ascontiguousarray(local_to_world, np.float32).reshape(-1, 16)
        check(self._lib.vgpu_render_set_camera_keys(self._h, ptr(m), len(m)))

    def render_pass(self, scrambles, iteration, want_sample=False):
        """One pass with host arrays (synchronous); updates self.film.  Returns (shadow rays, sample)."""
        scr = np.ascontiguousarray(scrambles, np.uint64)
        samp = np.zeros((self.height, self.width, 3), np.float32) if want_sample else None
        n = C.c_uint64(0)
        check(self._lib.vgpu_render_pass(self._h, 0, ptr(scr), iteration, ptr(self.film), ptr(samp), None, C.byref(n)))
        return n.value, samp

    def render_pass_device(self, d_scrambles, iteration, d_film, d_sample=0, stream=0, graph=False):
        """One pass on device buffers (ints: data pointers), enqueued on `stream`; graph=True replays the
        pass as a HIP graph."""
        flags = _lib.VGPU_TRACE_DEVICE_PTRS | (_lib.VGPU_GRAPH if graph else 0)
        check(self._lib.vgpu_render_pass(self._h, flags, C.c_void_p(d_scrambles), iteration,
                                         C.c_void_p(d_film), C.c_void_p(d_sample or None), C.c_void_p(stream or None),
                                         None))
