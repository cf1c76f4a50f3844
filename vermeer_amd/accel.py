"""Host-side mirror of the reference's traversal plugin interface.

The names follow the reference so that code (and tests) read alike:

  build_accel(boxes, centroids, indxs, leaf_max)   qbvh.BuildAccel      (qbvh/build.go:293)
  build_accel_gpu(...)                             qbvh.BuildAccel, built on the GPU
  build_accel_motion(...)                          qbvh.BuildAccelMotion (qbvh/motionbuild.go:108)
  PolyMesh                                         builtin/geom/polymesh/polymesh.go:17-60
  Scene.add_geom / pre_render / trace              builtin/scene/scene.go:119-133, :30-58 and
                                                   core.TraceProbe (core/trace.go:26-35), batched

All work happens in libvermeer_gpu.so (C++ host builder + HIP kernels);
this module only marshals numpy arrays or device tensors across the C ABI.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import HIT_DTYPE, MNODE_DTYPE, NODE_DTYPE, RAY_DTYPE, check, ptr


def _f32(a, shape_last):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != shape_last:
        raise ValueError(f"expected an (n, {shape_last}) float32 array, got {a.shape}")
    return a


def build_accel(boxes, centroids, indxs, leaf_max=16, nthreads=0):
    """qbvh.BuildAccel: returns (nodes, bounds).

    boxes (n,6) [min xyz, max xyz], centroids (n,3) and indxs (n,) are
    permuted IN PLACE exactly as the reference permutes its slices; they must
    be C-contiguous float32 / float32 / int32 numpy arrays.
    """
    n = len(indxs)
    for a, dt in ((boxes, np.float32), (centroids, np.float32), (indxs, np.int32)):
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
            raise TypeError("build_accel permutes its inputs in place: pass C-contiguous numpy arrays "
                            "(float32 boxes/centroids, int32 indxs)")
    lib = _lib.load()
    nodes = np.zeros(max(n, 1), dtype=NODE_DTYPE)
    nn = C.c_int32(0)
    bounds = np.zeros(6, dtype=np.float32)
    check(lib.vgpu_qbvh_build(ptr(boxes), ptr(centroids), ptr(indxs), n, leaf_max, ptr(nodes), len(nodes),
                              C.byref(nn), ptr(bounds), nthreads))
    return nodes[: nn.value].copy(), bounds


def build_accel_gpu(boxes, centroids, indxs, leaf_max=16, stream=None):
    """qbvh.BuildAccel computed on the current GPU (vgpu_qbvh_build_gpu): the same (nodes, bounds)
    and the same in-place permutation as build_accel.  stream: a hipStream_t as an int, or None."""
    n = len(indxs)
    for a, dt in ((boxes, np.float32), (centroids, np.float32), (indxs, np.int32)):
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
            raise TypeError("build_accel_gpu permutes its inputs in place: pass C-contiguous numpy arrays "
                            "(float32 boxes/centroids, int32 indxs)")
    lib = _lib.load()
    nodes = np.zeros(max(n, 1), dtype=NODE_DTYPE)
    nn = C.c_int32(0)
    bounds = np.zeros(6, dtype=np.float32)
    check(lib.vgpu_qbvh_build_gpu(ptr(boxes), ptr(centroids), ptr(indxs), n, leaf_max, ptr(nodes), len(nodes),
                                  C.byref(nn), ptr(bounds), C.c_void_p(stream or 0)))
    return nodes[: nn.value].copy(), bounds


def build_accel_motion(boxes, centroids, indxs, leaf_max=16, nthreads=0):
    """qbvh.BuildAccelMotion: returns the topology nodes (MNODE_DTYPE)."""
    n = len(indxs)
    for a, dt in ((boxes, np.float32), (centroids, np.float32), (indxs, np.int32)):
        if not (isinstance(a, np.ndarray) and a.dtype == dt and a.flags.c_contiguous):
            raise TypeError("build_accel_motion permutes its inputs in place: pass C-contiguous numpy arrays")
    lib = _lib.load()
    nodes = np.zeros(max(n, 1), dtype=MNODE_DTYPE)
    nn = C.c_int32(0)
    check(lib.vgpu_qbvh_build_motion(ptr(boxes), ptr(centroids), ptr(indxs), n, leaf_max, ptr(nodes), len(nodes),
                                     C.byref(nn), nthreads))
    return nodes[: nn.value].copy()


class PolyMesh:
    """A triangle mesh: verts (nkeys, nverts, 3) or (nverts, 3), face_idx (ntris, 3).

    Mirrors PolyMesh's Verts (param.PointArray with MotionKeys), FaceIdx and
    RayBias fields (builtin/geom/polymesh/polymesh.go:17-60).
    """

    def __init__(self, verts, face_idx, ray_bias=0.0, motion_reference_indexing=False, name=""):
        v = np.ascontiguousarray(verts, dtype=np.float32)
        if v.ndim == 2:
            v = v[None]
        if v.ndim != 3 or v.shape[2] != 3:
            raise ValueError("verts must be (nverts,3) or (nkeys,nverts,3)")
        self.verts = v
        self.face_idx = np.ascontiguousarray(face_idx, dtype=np.uint32).reshape(-1, 3)
        self.ray_bias = float(ray_bias)
        self.motion_reference_indexing = bool(motion_reference_indexing)
        self.name = name

    @property
    def motion_keys(self):
        return self.verts.shape[0]

    @property
    def ntris(self):
        return self.face_idx.shape[0]


class Scene:
    """The scene plug-in (core.Scene, core/scene.go:8-24) backed by the GPU.

    add_geom() then pre_render() (builds every accel on the host exactly as
    the reference would and uploads it); trace() runs batched probes.
    """

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.vgpu_scene_create(C.byref(h), device))
        self._h = h
        self.device = device
        self.geoms = []
        self.committed = False

    @classmethod
    def adopt(cls, handle, device=0):
        """A Scene owning a committed vgpu_scene the library created (vgpu_vnf_build)."""
        sc = cls.__new__(cls)
        sc._lib = _lib.load()
        sc._h = handle
        sc.device = device
        sc.geoms = []
        sc.committed = True
        return sc

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vgpu_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_geom(self, mesh):
        g = C.c_int32(-1)
        flags = _lib.VGPU_MESH_MOTION_REFERENCE_INDEXING if mesh.motion_reference_indexing else 0
        check(self._lib.vgpu_scene_add_mesh(self._h, ptr(mesh.verts), mesh.verts.shape[1], mesh.verts.shape[0],
                                            ptr(mesh.face_idx), mesh.ntris, mesh.ray_bias, flags, C.byref(g)))
        self.geoms.append(mesh)
        return g.value

    def add_prebuilt(self, verts, idxp_sorted, nodes, bounds, ray_bias=0.0):
        """Static mesh with an accel built by the reference's own builder (the cgo path)."""
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
        ix = np.ascontiguousarray(idxp_sorted, dtype=np.uint32).reshape(-1, 3)
        nd = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        b = np.ascontiguousarray(bounds, dtype=np.float32).reshape(6)
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_mesh_prebuilt(self._h, ptr(v), len(v), ptr(ix), len(ix), ptr(nd), len(nd),
                                                     ptr(b), float(ray_bias), C.byref(g)))
        self.geoms.append(None)
        return g.value

    def add_motion_prebuilt(self, verts, idxp, accel_idx, nodes, boxes, ray_bias=0.0, reference_indexing=False):
        """Moving mesh with an accel built by the reference's own code (the cgo path): verts (nkeys, nverts, 3),
        idxp (ntris, 3) in face order, accel_idx (ntris,) faces in leaf order, nodes (MNODE_DTYPE or an
        (nnodes, 8) int32 array: axis0..2, children[4], pad), boxes (nkeys, nnodes, 24)."""
        v = np.ascontiguousarray(verts, dtype=np.float32)
        if v.ndim != 3 or v.shape[2] != 3 or v.shape[0] < 2:
            raise ValueError("verts must be (nkeys >= 2, nverts, 3)")
        ix = np.ascontiguousarray(idxp, dtype=np.uint32).reshape(-1, 3)
        ai = np.ascontiguousarray(accel_idx, dtype=np.int32).reshape(-1)
        nd = np.ascontiguousarray(nodes)
        if nd.dtype != MNODE_DTYPE:
            nd = np.ascontiguousarray(nd, dtype=np.int32).reshape(-1, 8).view(MNODE_DTYPE).reshape(-1)
        bx = np.ascontiguousarray(boxes, dtype=np.float32)
        if bx.shape != (v.shape[0], len(nd), 24):
            raise ValueError(f"boxes must be (nkeys, nnodes, 24), got {bx.shape}")
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_motion_mesh_prebuilt(
            self._h, ptr(v), v.shape[1], v.shape[0], ptr(ix), len(ix), ptr(ai), ptr(nd), len(nd), ptr(bx),
            float(ray_bias), 1 if reference_indexing else 0, C.byref(g)))
        self.geoms.append(None)
        return g.value

    def add_sphere(self, centre, radius, material):
        """The analytic sphere a SphereLight inserts (builtin/geom/sphere) with its shader (a material id)."""
        c = np.ascontiguousarray(centre, dtype=np.float32).reshape(3)
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_sphere(self._h, ptr(c), float(radius), int(material), C.byref(g)))
        self.geoms.append(None)
        return g.value

    def set_transform(self, geom, matrices):
        """A mesh's Transform keys, (nkeys, 16) column-major (vgpu_scene_set_transform); geom: insertion index."""
        m = np.ascontiguousarray(matrices, np.float32).reshape(-1, 16)
        check(self._lib.vgpu_scene_set_transform(self._h, int(geom), ptr(m), len(m)))

    def set_hidden(self, geom, hidden=True):
        """Leaves a geom out of the top level: traced only through instances and procs holding it."""
        check(self._lib.vgpu_scene_set_hidden(self._h, int(geom), 1 if hidden else 0))

    def add_instance(self, geom, matrices, bmin, bmax):
        """GeomInstance of geom (insertion index) under transform keys (nkeys, 16); returns its insertion index."""
        m = np.ascontiguousarray(matrices, np.float32).reshape(-1, 16)
        lo, hi = np.ascontiguousarray(bmin, np.float32), np.ascontiguousarray(bmax, np.float32)
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_instance(self._h, int(geom), ptr(m), len(m), ptr(lo), ptr(hi), C.byref(g)))
        self.geoms.append(None)
        return g.value

    def add_proc(self, geoms, matrices, bmin, bmax):
        """Proc over geoms (insertion indices, in order) under transform keys (nkeys, 16) or None (identity)."""
        ch = np.ascontiguousarray(geoms, np.int32)
        m = None if matrices is None else np.ascontiguousarray(matrices, np.float32).reshape(-1, 16)
        lo, hi = np.ascontiguousarray(bmin, np.float32), np.ascontiguousarray(bmax, np.float32)
        g = C.c_int32(-1)
        check(self._lib.vgpu_scene_add_proc(self._h, ptr(ch), len(ch), ptr(m), 0 if m is None else len(m), ptr(lo),
                                            ptr(hi), C.byref(g)))
        self.geoms.append(None)
        return g.value

    def add_transforms(self, desc):
        """The transforms of a scene description: per mesh "transform" (nkeys, 16) and "hidden", then
        desc["instances"] ({geom, matrices, bmin, bmax}) and desc["procs"] ({geoms, matrices or None, bmin,
        bmax}), added after the meshes in that order."""
        for gi, mesh in enumerate(desc["meshes"]):
            if mesh.get("transform") is not None:
                self.set_transform(gi, mesh["transform"])
            if mesh.get("hidden"):
                self.set_hidden(gi)
        for it in desc.get("instances", []):
            self.add_instance(it["geom"], it["matrices"], it["bmin"], it["bmax"])
        for pr in desc.get("procs", []):
            self.add_proc(pr["geoms"], pr.get("matrices"), pr["bmin"], pr["bmax"])

    def pre_render(self):
        check(self._lib.vgpu_scene_commit(self._h))
        self.committed = True

    def info(self):
        ng = C.c_int32(0)
        nt = C.c_int32(0)
        n = len(self.geoms)
        order = np.zeros(n, np.int32)
        nodes = np.zeros(n, np.int32)
        tris = np.zeros(n, np.int32)
        check(self._lib.vgpu_scene_info(self._h, C.byref(ng), C.byref(nt), ptr(order), ptr(nodes), ptr(tris)))
        return {"ngeoms": ng.value, "top_nodes": nt.value, "geom_order": order, "geom_nodes": nodes, "geom_tris": tris}

    def mesh_accel(self, slot):
        inf = self.info()
        nodes = np.zeros(inf["geom_nodes"][slot], NODE_DTYPE)
        ntris = inf["geom_tris"][slot]
        idxp = np.zeros((ntris, 3), np.uint32)
        perm = np.zeros(ntris, np.int32)
        check(self._lib.vgpu_scene_mesh_accel(self._h, slot, ptr(nodes), ptr(idxp), ptr(perm)))
        return nodes, idxp, perm

    def trace(self, rays, shadow=False, want_geoms=True, want_counters=False):
        """Host arrays in, host arrays out (synchronous).

        rays: RAY_DTYPE array.  Returns (hits, geoms, counters) where hits is a
        HIT_DTYPE array, geoms an int32 array (or None), counters (n,4) int32
        (nodes, leaves, triangle tests, max stack depth) or None.
        """
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, HIT_DTYPE)
        geoms = np.zeros(n, np.int32) if want_geoms else None
        cnt = np.zeros((n, 4), np.int32) if want_counters else None
        flags = _lib.VGPU_TRACE_SHADOW if shadow else 0
        check(self._lib.vgpu_trace(self._h, flags, ptr(rays), n, ptr(hits), ptr(geoms), ptr(cnt), None))
        return hits, geoms, cnt

    def trace_device(self, d_rays, n, d_hits, shadow=False, d_geoms=0, d_counters=0, stream=0):
        """Device pointers (ints) in, enqueued on `stream` (an int hipStream_t); returns at once."""
        flags = _lib.VGPU_TRACE_DEVICE_PTRS | (_lib.VGPU_TRACE_SHADOW if shadow else 0)
        check(self._lib.vgpu_trace(self._h, flags, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                   C.c_void_p(d_geoms or None), C.c_void_p(d_counters or None),
                                   C.c_void_p(stream or None)))

    def status(self, stream=0):
        check(self._lib.vgpu_scene_status(self._h, C.c_void_p(stream or None)))

    def set_trace_config(self, block=256, lds_stack=16, waves_per_cu=16, sort=0, refill=0):
        cfg = _lib.TraceConfig(block, lds_stack, waves_per_cu, sort, refill)
        check(self._lib.vgpu_scene_set_trace_config(self._h, C.byref(cfg)))

    def set_leaf_size(self, leaf_max):
        """leafMax of the static meshes' BVH build (before pre_render; 16 = the reference's)."""
        check(self._lib.vgpu_scene_set_leaf_size(self._h, int(leaf_max)))

    def set_timing(self, enable=True):
        check(self._lib.vgpu_scene_set_timing(self._h, int(bool(enable))))

    def timing(self):
        """(traversal ms, sort ms, traversals) summed since the last call (timing must be enabled)."""
        t, so, k = C.c_float(), C.c_float(), C.c_int32()
        check(self._lib.vgpu_scene_timing(self._h, C.byref(t), C.byref(so), C.byref(k)))
        return float(t.value), float(so.value), int(k.value)

    def trace_config(self):
        cfg = _lib.TraceConfig()
        check(self._lib.vgpu_scene_get_trace_config(self._h, C.byref(cfg)))
        return cfg.as_dict()

    def autotune(self, d_rays, n, d_hits, shadow=False, stream=0):
        cfg = _lib.TraceConfig()
        ms = C.c_float(0)
        flags = _lib.VGPU_TRACE_SHADOW if shadow else 0
        check(self._lib.vgpu_scene_autotune(self._h, flags, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                            C.c_void_p(stream or None), C.byref(cfg), C.byref(ms)))
        return cfg.as_dict(), ms.value
