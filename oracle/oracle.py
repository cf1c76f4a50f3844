"""ORACLE — test infrastructure only.

ctypes wrapper of liboracle.so, the CPU restatement of the reference's
traversal path (oracle_accel.cpp).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module, and only as the checker
or the timed CPU baseline — never as part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("boxes", "<f4", 24), ("axis", "<u4", 3), ("children", "<i4", 4), ("parent", "<i4")])
RAY_DTYPE = np.dtype([("P", "<f4", 3), ("tmax", "<f4"), ("D", "<f4", 3), ("time", "<f4")])
OHIT_DTYPE = np.dtype([("geom", "<i4"), ("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

_lib = None


def build():
    """Compiles liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.c_void_p
    sig = {
        "orc_build_qbvh": (C.c_int, [P, P, P, C.c_int32, C.c_int32, P, C.c_int32, P]),
        "orc_intersect_boxes": (None, [P, P, P, P, P]),
        "orc_ray_setup": (None, [P, P, P, P]),
        "orc_scene_new": (P, []),
        "orc_scene_free": (None, [P]),
        "orc_scene_add_mesh": (C.c_int, [P, P, C.c_int32, C.c_int32, P, C.c_int32, C.c_float, C.c_int32]),
        "orc_scene_add_sphere": (C.c_int, [P, P, C.c_float, C.c_int32]),
        "orc_scene_commit": (C.c_int, [P]),
        "orc_scene_set_transform": (None, [P, C.c_int32, P, C.c_int32]),
        "orc_scene_set_hidden": (None, [P, C.c_int32, C.c_int32]),
        "orc_scene_add_instance": (C.c_int, [P, C.c_int32, P, C.c_int32, P, P]),
        "orc_scene_add_proc": (C.c_int, [P, P, C.c_int32, P, C.c_int32, P, P]),
        "orc_scene_set_leaf_size": (None, [P, C.c_int32]),
        "orc_scene_info": (C.c_int, [P, P, P, P, P, P]),
        "orc_scene_mesh_accel": (C.c_int, [P, C.c_int32, P, P, P]),
        "orc_scene_mesh_motion_accel": (C.c_int, [P, C.c_int32, P, P, P]),
        "orc_scene_top_accel": (C.c_int, [P, P]),
        "orc_scene_top_motion_accel": (C.c_int, [P, P, P]),
        "orc_trace": (C.c_int, [P, P, C.c_int64, C.c_int32, P, P, C.c_int32]),
        "orc_brute": (C.c_int, [P, P, C.c_int64, P, C.c_int32]),
        "orc_set_normalize_ieee": (None, [C.c_int32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def set_normalize_ieee(on):
    """Vec3Normalize estimate: False = RSQRTSS (the reference on this CPU), True = IEEE 1/sqrt."""
    load().orc_set_normalize_ieee(1 if on else 0)


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def build_qbvh(boxes, centroids, idx, leaf_max=16):
    """Restated qbvh.BuildAccel; permutes boxes/centroids/idx in place. Returns (nodes, bounds)."""
    lib = load()
    n = len(idx)
    nodes = np.zeros(max(n, 1), NODE_DTYPE)
    bounds = np.zeros(6, np.float32)
    rc = lib.orc_build_qbvh(_p(boxes), _p(centroids), _p(idx), n, leaf_max, _p(nodes), len(nodes), _p(bounds))
    if rc < 0:
        raise RuntimeError(f"oracle build failed ({rc})")
    return nodes[:rc].copy(), bounds


def intersect_boxes(P, D, boxes24):
    lib = load()
    P = np.ascontiguousarray(P, np.float32)
    D = np.ascontiguousarray(D, np.float32)
    b = np.ascontiguousarray(boxes24, np.float32)
    hits = np.zeros(4, np.int32)
    t = np.zeros(4, np.float32)
    lib.orc_intersect_boxes(_p(P), _p(D), _p(b), _p(hits), _p(t))
    return hits, t


def ray_setup(D):
    lib = load()
    D = np.ascontiguousarray(D, np.float32)
    S = np.zeros(3, np.float32)
    Dinv = np.zeros(3, np.float32)
    K = np.zeros(3, np.int32)
    lib.orc_ray_setup(_p(D), _p(S), _p(Dinv), _p(K))
    return S, Dinv, K


class OracleScene:
    """Scene of PolyMeshes, built and traced exactly as the reference does (on the CPU)."""

    def __init__(self):
        self.lib = load()
        self.h = self.lib.orc_scene_new()
        self.meshes = []

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_scene_free(self.h)
            self.h = None

    def add_mesh(self, verts, idx, ray_bias=0.0, motion_reference_indexing=True):
        v = np.ascontiguousarray(verts, np.float32)
        if v.ndim == 2:
            v = v[None]
        ix = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
        g = self.lib.orc_scene_add_mesh(self.h, _p(v), v.shape[1], v.shape[0], _p(ix), len(ix), float(ray_bias),
                                        1 if motion_reference_indexing else 0)
        self.meshes.append((v, ix))
        return g

    def add_sphere(self, centre, radius, material=0):
        """A SphereLight's analytic sphere (builtin/geom/sphere); hits report prim 0, u = v = 0."""
        c = np.ascontiguousarray(centre, np.float32)
        g = self.lib.orc_scene_add_sphere(self.h, _p(c), float(radius), int(material))
        self.meshes.append(None)
        return g

    def set_transform(self, geom, mats):
        """A mesh's Transform keys: (nkeys, 16) column-major matrices (builtin/geom/polymesh/trace.go:20-74)."""
        m = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        self.lib.orc_scene_set_transform(self.h, int(geom), _p(m), len(m))

    def set_hidden(self, geom, hidden=True):
        """Leaves a geom out of the scene's top level (reached only through instances or procs)."""
        self.lib.orc_scene_set_hidden(self.h, int(geom), 1 if hidden else 0)

    def add_instance(self, proto, mats, bmin, bmax):
        """GeomInstance of geom `proto` under transform keys (nkeys, 16); returns its insertion index."""
        m = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        lo, hi = np.ascontiguousarray(bmin, np.float32), np.ascontiguousarray(bmax, np.float32)
        return self.lib.orc_scene_add_instance(self.h, int(proto), _p(m), len(m), _p(lo), _p(hi))

    def add_proc(self, children, mats, bmin, bmax):
        """Proc over geoms `children` (in order) under transform keys (nkeys, 16) or None (identity)."""
        ch = np.ascontiguousarray(children, np.int32)
        m = np.ascontiguousarray(mats if mats is not None else np.zeros((0, 16)), np.float32).reshape(-1, 16)
        lo, hi = np.ascontiguousarray(bmin, np.float32), np.ascontiguousarray(bmax, np.float32)
        return self.lib.orc_scene_add_proc(self.h, _p(ch), len(ch), _p(m), len(m), _p(lo), _p(hi))

    def commit(self, leaf_size=16):
        self.lib.orc_scene_set_leaf_size(self.h, int(leaf_size))
        if self.lib.orc_scene_commit(self.h) != 0:
            raise RuntimeError("oracle scene build failed")

    def info(self):
        n = len(self.meshes)
        ntop = np.zeros(1, np.int32)
        order, nodes, tris, keys = (np.zeros(n, np.int32) for _ in range(4))
        self.lib.orc_scene_info(self.h, _p(ntop), _p(order), _p(nodes), _p(tris), _p(keys))
        return {"top_nodes": int(ntop[0]), "geom_order": order, "geom_nodes": nodes, "geom_tris": tris, "geom_keys": keys}

    def mesh_accel(self, slot):
        inf = self.info()
        nodes = np.zeros(inf["geom_nodes"][slot], NODE_DTYPE)
        ntris = inf["geom_tris"][slot]
        idxp = np.zeros((ntris, 3), np.uint32)
        perm = np.zeros(ntris, np.int32)
        if self.lib.orc_scene_mesh_accel(self.h, slot, _p(nodes), _p(idxp), _p(perm)) != 0:
            raise RuntimeError("not a static mesh")
        return nodes, idxp, perm

    def mesh_motion_accel(self, slot):
        inf = self.info()
        nn, ntris, nk = inf["geom_nodes"][slot], inf["geom_tris"][slot], inf["geom_keys"][slot]
        topo = np.zeros((nn, 8), np.int32)
        boxes = np.zeros((nk, nn, 24), np.float32)
        perm = np.zeros(ntris, np.int32)
        if self.lib.orc_scene_mesh_motion_accel(self.h, slot, _p(topo), _p(boxes), _p(perm)) != 0:
            raise RuntimeError("not a motion mesh")
        return topo, boxes, perm

    def top_accel(self):
        nodes = np.zeros(self.info()["top_nodes"], NODE_DTYPE)
        if self.lib.orc_scene_top_accel(self.h, _p(nodes)) != 0:
            raise RuntimeError("top level is a motion QBVH")
        return nodes

    def trace(self, rays, shadow=False, counters=False, nthreads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, OHIT_DTYPE)
        cnt = np.zeros((n, 4), np.int32) if counters else None
        self.lib.orc_trace(self.h, _p(rays), n, 1 if shadow else 0, _p(hits), _p(cnt), nthreads)
        return hits, cnt

    def brute(self, rays, nthreads=0):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), OHIT_DTYPE)
        self.lib.orc_brute(self.h, _p(rays), len(rays), _p(hits), nthreads)
        return hits


# ------------------------------------------------------------------ renderer

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


def _render_sigs(lib):
    P = C.c_void_p
    sig = {
        "orc_camera_lookat": (None, [P, P, P, C.c_float, C.c_float, P, P]),
        "orc_vdc": (C.c_double, [C.c_uint64, C.c_uint64]),
        "orc_sobol": (C.c_double, [C.c_uint64, C.c_uint64]),
        "orc_raster_xy": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P]),
        "orc_raster_tables": (None, [C.c_int32, P, P]),
        "orc_smits": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_float]),
        "orc_spectrum_to_rgb": (None, [P, C.c_float, P]),
        "orc_cie_tables": (None, [P, P, P]),
        "orc_scene_add_shaded_mesh": (C.c_int, [P, P, C.c_int32, P, C.c_int32, P, P, P, C.c_int32]),
        "orc_camera_lookat_keys": (C.c_int, [P, C.c_int32, P, C.c_int32, P, P, C.c_int32, P]),
        "orc_render_set_camera_keys": (None, [P, P, C.c_int32]),
        "orc_decomp_lerp_matrix": (None, [P, P, C.c_float, P]),
        "orc_scene_add_shaded_motion_mesh": (C.c_int, [P, P, C.c_int32, C.c_int32, P, C.c_int32, P, C.c_int32, P, P,
                                                       C.c_int32, C.c_int32]),
        "orc_render_new": (P, [P, C.c_int32, C.c_int32, P, C.c_float, C.c_float, C.c_float, P, C.c_int32]),
        "orc_render_free": (None, [P]),
        "orc_render_add_light": (None, [P, C.c_int32, P, P, P, C.c_float, C.c_int32, C.c_int32, C.c_int32]),
        "orc_render_set_filter": (None, [P, C.c_int32, C.c_float, C.c_int32, C.c_float]),
        "orc_render_set_lens": (None, [P, C.c_float]),
        "orc_render_set_path_depth": (None, [P, C.c_int32]),
        "orc_scene_mesh_set_normals": (None, [P, C.c_int32, P, C.c_int32, P]),
        "orc_scene_mesh_set_ray_bias": (None, [P, C.c_int32, C.c_float]),
        "orc_uniform_disk": (None, [C.c_float, C.c_float, C.c_float, P]),
        "orc_render_warp_sample": (None, [P, C.c_double, C.c_double, P]),
        "orc_render_filter_tables": (C.c_int32, [P, P]),
        "orc_fresnel_conductor": (C.c_float, [C.c_float, C.c_float, C.c_float]),
        "orc_render_pass": (C.c_int64, [P, P, C.c_uint32, P, P, C.c_int32]),
        "orc_fresnel_dielectric": (C.c_float, [C.c_float, C.c_float]),
        "orc_ggx_d": (C.c_float, [P, C.c_float]),
        "orc_oren_nayar_eval": (None, [C.c_float, P, P, C.c_float, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def camera_lookat(frm, to, up, fov_deg, focal=12.0):
    """Camera matrix (column-major 4x4) and TanThetaFocal as the reference computes them."""
    lib = load()
    _render_sigs(lib)
    M = np.zeros(16, np.float32)
    ttf = np.zeros(1, np.float32)
    f = np.asarray(frm, np.float32)
    t = np.asarray(to, np.float32)
    u = np.asarray(up, np.float32)
    lib.orc_camera_lookat(_p(f), _p(t), _p(u), float(fov_deg), float(focal), _p(M), _p(ttf))
    return M, float(ttf[0])


def camera_lookat_keys(from_keys, to_keys, up, roll=None):
    """Camera.calcLookatMatrices (camera.go:109-193): (nkeys, 16) LocalToWorld matrices for keyed From /
    To (and Roll, radians)."""
    lib = load(); _render_sigs(lib)
    f = np.ascontiguousarray(from_keys, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(to_keys, np.float32).reshape(-1, 3)
    u = np.ascontiguousarray(up, np.float32)
    r = None if roll is None else np.ascontiguousarray(roll, np.float32).reshape(-1)
    out = np.zeros((max(len(f), len(t)), 16), np.float32)
    n = lib.orc_camera_lookat_keys(_p(f), len(f), _p(t), len(t), _p(u), _p(r), 0 if r is None else len(r), _p(out))
    return out[:n]


def decomp_lerp_matrix(a, b, t):
    """TransformDecompToMatrix4(TransformDecompLerp(decomp(a), decomp(b), t)) (animdecomp.go:20-80)."""
    lib = load(); _render_sigs(lib)
    a = np.ascontiguousarray(a, np.float32).reshape(16)
    b = np.ascontiguousarray(b, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    lib.orc_decomp_lerp_matrix(_p(a), _p(b), float(t), _p(out))
    return out


def vdc(i, s=0):
    lib = load(); _render_sigs(lib)
    return lib.orc_vdc(i, s)


def sobol(i, s=0):
    lib = load(); _render_sigs(lib)
    return lib.orc_sobol(i, s)


def raster_xy(m, frame, px, py):
    lib = load(); _render_sigs(lib)
    idx = np.zeros(1, np.uint64)
    rx = np.zeros(1, np.float64)
    ry = np.zeros(1, np.float64)
    lib.orc_raster_xy(m, frame, px, py, _p(idx), _p(rx), _p(ry))
    return int(idx[0]), float(rx[0]), float(ry[0])


def raster_tables(m):
    lib = load(); _render_sigs(lib)
    fwd = np.zeros(52, np.uint64)
    inv = np.zeros(52, np.uint64)
    lib.orc_raster_tables(m, _p(fwd), _p(inv))
    return fwd, inv


def smits(r, g, b, lam):
    lib = load(); _render_sigs(lib)
    return lib.orc_smits(r, g, b, lam)


def spectrum_to_rgb(C4, lam):
    lib = load(); _render_sigs(lib)
    c = np.ascontiguousarray(C4, np.float32)
    out = np.zeros(3, np.float32)
    lib.orc_spectrum_to_rgb(_p(c), float(lam), _p(out))
    return out


def cie_tables():
    lib = load(); _render_sigs(lib)
    x, y, z = (np.zeros(95, np.float32) for _ in range(3))
    lib.orc_cie_tables(_p(x), _p(y), _p(z))
    return x, y, z


class OracleRenderer:
    """The reference's render loop (reference semantics) on the CPU.

    scene_desc: dict from vermeer_amd.scenes (meshes with shader tables, materials, tri lights, camera).
    """

    def __init__(self, desc, width, height, M, tan_theta_focal, focal=12.0, aspect=None):
        self.lib = load()
        _render_sigs(self.lib)
        self.scene = OracleScene()
        self.w, self.h = width, height
        self._keep = []
        for mesh in desc["meshes"]:
            if mesh.get("instance") is not None:   # an instance or proc in scene order (a .vnf's AddNode order)
                it = mesh["instance"]
                self.scene.add_instance(it["geom"], it["matrices"], it["bmin"], it["bmax"])
                self.scene.meshes.append(None)
                continue
            if mesh.get("proc") is not None:
                pr = mesh["proc"]
                self.scene.add_proc(pr["geoms"], pr.get("matrices"), pr["bmin"], pr["bmax"])
                self.scene.meshes.append(None)
                continue
            if mesh.get("sphere"):
                c = np.ascontiguousarray(mesh["centre"], np.float32)
                self._keep.append(c)
                self.lib.orc_scene_add_sphere(self.scene.h, _p(c), float(mesh["radius"]), int(mesh["material"]))
                self.scene.meshes.append(None)
                continue
            if int(mesh.get("keys", 1)) > 1:
                self._add_motion_mesh(mesh)
                continue
            v = np.ascontiguousarray(mesh["verts"], np.float32)
            ix = np.ascontiguousarray(mesh["idx"], np.uint32)
            nrm = mesh.get("normals")
            nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
            sidx = mesh.get("shaderidx")
            sidx = None if sidx is None else np.ascontiguousarray(sidx, np.uint8)
            shaders = np.ascontiguousarray(mesh["shaders"], np.int32)
            self._keep += [v, ix, nrm, sidx, shaders]
            g = self.lib.orc_scene_add_shaded_mesh(self.scene.h, _p(v), len(v), _p(ix), len(ix), _p(nrm), _p(sidx),
                                                   _p(shaders), len(shaders))
            if mesh.get("normal_list") is not None:
                # Normals with their own index per face corner (PolyMesh.normalidx)
                nl = np.ascontiguousarray(mesh["normal_list"], np.float32).reshape(-1, 3)
                ni = np.ascontiguousarray(mesh["normal_idx"], np.uint32)
                self._keep += [nl, ni]
                self.lib.orc_scene_mesh_set_normals(self.scene.h, g, _p(nl), len(nl), _p(ni))
            if mesh.get("ray_bias"):
                self.lib.orc_scene_mesh_set_ray_bias(self.scene.h, g, float(mesh["ray_bias"]))
            self.scene.meshes.append((v, ix))
        for gi, mesh in enumerate(desc["meshes"]):
            if mesh.get("transform") is not None:
                self.scene.set_transform(gi, mesh["transform"])
            if mesh.get("hidden"):
                self.scene.set_hidden(gi)
        for it in desc.get("instances", []):
            self.scene.add_instance(it["geom"], it["matrices"], it["bmin"], it["bmax"])
        for pr in desc.get("procs", []):
            self.scene.add_proc(pr["geoms"], pr.get("matrices"), pr["bmin"], pr["bmax"])
        self.scene.commit()
        mats = np.ascontiguousarray(desc["materials"], MATERIAL_DTYPE)
        Mx = np.ascontiguousarray(M, np.float32)
        asp = float(width) / float(height) if aspect is None else float(aspect)
        self.h_r = self.lib.orc_render_new(self.scene.h, width, height, _p(Mx), float(tan_theta_focal),
                                           asp, float(focal), _p(mats), len(mats))
        types = {"tri": 0, "disk": 1, "sphere": 2}
        for L in desc["lights"]:
            kind = L.get("type", "tri")
            keys = {"tri": ("P0", "P1", "P2"), "disk": ("P", "Up", "LookAt"), "sphere": ("P", "P", "P")}[kind]
            p = [np.ascontiguousarray(L[k], np.float32) for k in keys]
            self.lib.orc_render_add_light(self.h_r, types[kind], _p(p[0]), _p(p[1]), _p(p[2]),
                                          float(L.get("radius", 0.0)), int(L["material"]), int(L["samples"]),
                                          int(L["geom"]))
        self.film = np.zeros((height, width, 3), np.float32)

    def _add_motion_mesh(self, mesh):
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
        shaders = np.ascontiguousarray(mesh["shaders"], np.int32)
        self._keep += [v, ix, nrm, sidx, shaders]
        g = self.lib.orc_scene_add_shaded_motion_mesh(self.scene.h, _p(v), v.shape[1], nk, _p(ix), len(ix), _p(nrm),
                                                      nrm_keys, _p(sidx), _p(shaders), len(shaders),
                                                      1 if mesh.get("reference_indexing") else 0)
        if mesh.get("ray_bias"):
            self.lib.orc_scene_mesh_set_ray_bias(self.scene.h, g, float(mesh["ray_bias"]))
        self.scene.meshes.append((v, ix))

    def __del__(self):
        if getattr(self, "h_r", None):
            self.lib.orc_render_free(self.h_r)
            self.h_r = None

    def set_lens(self, radius):
        """The camera node's Radius: a thin lens when > 0 (camera.go:265-283)."""
        self.lib.orc_render_set_lens(self.h_r, float(radius))

    def set_path_depth(self, depth):
        """Path mode (no reference counterpart): paths of up to `depth` segments; 0 keeps the
        reference's shading."""
        self.lib.orc_render_set_path_depth(self.h_r, int(depth))

    def set_filter(self, kind="box", width=0.0, res=0, peak=0.0):
        t = {"box": 0, "gaussian": 1, "airy": 2}[kind]
        self.lib.orc_render_set_filter(self.h_r, t, float(width), int(res), float(peak))

    def warp_sample(self, r0, r1):
        uv = np.zeros(2, np.float64)
        self.lib.orc_render_warp_sample(self.h_r, float(r0), float(r1), _p(uv))
        return float(uv[0]), float(uv[1])

    def filter_tables(self):
        """(cdfV, cdfVU) of the current filter, or None for the box filter."""
        n = self.lib.orc_render_filter_tables(self.h_r, None)
        if n == 0:
            return None
        out = np.zeros(n + n * n, np.float64)
        self.lib.orc_render_filter_tables(self.h_r, _p(out))
        return out[:n], out[n:].reshape(n, n)

    def set_camera_keys(self, mats):
        m = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        self._keep.append(m)
        self.lib.orc_render_set_camera_keys(self.h_r, _p(m), len(m))

    def light_samples(self, li, P, Ng, lam, s0, s1, I0, count, n=1):
        """Light.SampleArea of light li at (P, Ng) for sample indices I0..I0+count-1, n each:
        (m, 5) float32 rows of (Ld xyz, Ldist, Pdf), accepted samples only."""
        _pin_sigs(self.lib)
        out = np.zeros((count * n, 5), np.float32)
        Pv, Nv = _v3(P), _v3(Ng)      # held until the call returns
        m = self.lib.orc_light_sample_range(self.h_r, li, _p(Pv), _p(Nv), float(lam), int(s0), int(s1),
                                            int(I0), int(count), int(n), _p(out))
        return out[:m]

    def eval_light_samples(self, li, P, N, dPdu, Rd, lam, roughness, scrambles, I0, count, nsamples):
        """EvaluateLightSamples of light li with ShaderStd's diffuse lobe: (count, 3) RGB per index."""
        _pin_sigs(self.lib)
        scr = np.ascontiguousarray(scrambles, np.uint64).reshape(2)
        out = np.zeros((count, 3), np.float32)
        vs = [_v3(P), _v3(N), _v3(dPdu), _v3(Rd)]      # held until the call returns
        self.lib.orc_eval_light_samples(self.h_r, li, _p(vs[0]), _p(vs[1]), _p(vs[2]), _p(vs[3]), float(lam),
                                        float(roughness), _p(scr), int(I0), int(count), int(nsamples), _p(out))
        return out

    def render_pass(self, scrambles, iteration, nthreads=0, want_sample=False):
        scr = np.ascontiguousarray(scrambles, np.uint64)
        samp = np.zeros((self.h, self.w, 3), np.float32) if want_sample else None
        n = self.lib.orc_render_pass(self.h_r, _p(scr), iteration, _p(self.film), _p(samp), nthreads)
        return n, samp


def fresnel_dielectric(eta, c):
    lib = load(); _render_sigs(lib)
    return lib.orc_fresnel_dielectric(eta, c)


def uniform_disk(radius, r0, r1):
    lib = load(); _render_sigs(lib)
    xy = np.zeros(2, np.float32)
    lib.orc_uniform_disk(float(radius), float(r0), float(r1), _p(xy))
    return float(xy[0]), float(xy[1])


def fresnel_conductor(r, g, c):
    lib = load(); _render_sigs(lib)
    return lib.orc_fresnel_conductor(r, g, c)


def ggx_d(m, alpha):
    lib = load(); _render_sigs(lib)
    mm = np.ascontiguousarray(m, np.float32)
    return lib.orc_ggx_d(_p(mm), alpha)


def oren_nayar_eval(lam, wi, wo, roughness):
    lib = load(); _render_sigs(lib)
    a = np.ascontiguousarray(wi, np.float32)
    b = np.ascontiguousarray(wo, np.float32)
    out = np.zeros(4, np.float32)
    lib.orc_oren_nayar_eval(lam, _p(a), _p(b), roughness, _p(out))
    return out


def _pin_sigs(lib):
    P = C.c_void_p
    sig = {
        "orc_sample_spherical_triangle_n": (None, [P, P, P, P, C.c_int64, P, P, P, P]),
        "orc_light_sample_range": (C.c_int64, [P, C.c_int32, P, P, C.c_float, C.c_uint64, C.c_uint64, C.c_uint64,
                                               C.c_int64, C.c_int32, P]),
        "orc_eval_light_samples": (None, [P, C.c_int32, P, P, P, P, C.c_float, C.c_float, P, C.c_uint64, C.c_int64,
                                          C.c_int32, P]),
        "orc_spectrum_product_rgb": (None, [P, P, C.c_float, P]),
        "orc_ggx_sample_n": (None, [C.c_float, P, C.c_int64, P, P, P]),
        "orc_ggx_pdf_eval_n": (None, [C.c_float, C.c_float, C.c_float, P, C.c_int64, P, P, P]),
        "orc_ggx_d_n": (None, [P, C.c_int64, C.c_float, P]),
        "orc_ggx_g1_n": (None, [P, P, C.c_int64, C.c_float, P]),
        "orc_spectrum_rgb_n": (None, [P, C.c_int64, P, P]),
        "orc_path_lobe_weight_n": (None, [C.c_int32, C.c_float, C.c_float, C.c_float, P, C.c_int64, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _v3(a):
    return np.ascontiguousarray(a, np.float32).reshape(3)


def sample_spherical_triangle(p0, p1, p2, p, r0, r1):
    """sampleSphericalTriangle (builtin/light/triangle.go:474-535) for arrays r0, r1: (dirs (n,3), pdf (n,))."""
    lib = _pin_sigs(load())
    r0 = np.ascontiguousarray(r0, np.float64)
    r1 = np.ascontiguousarray(r1, np.float64)
    n = len(r0)
    dirs = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float64)
    a, b, c, o = _v3(p0), _v3(p1), _v3(p2), _v3(p)
    lib.orc_sample_spherical_triangle_n(_p(a), _p(b), _p(c), _p(o), n, _p(r0), _p(r1), _p(dirs), _p(pdf))
    return dirs, pdf


def spectrum_product_rgb(a, b, lam):
    """ToRGB(FromRGB(a) * FromRGB(b)) at hero wavelength lam."""
    lib = _pin_sigs(load())
    out = np.zeros(3, np.float32)
    x, y = _v3(a), _v3(b)
    lib.orc_spectrum_product_rgb(_p(x), _p(y), float(lam), _p(out))
    return out


def ggx_sample(roughness, omega_r, r0, r1):
    lib = _pin_sigs(load())
    r0 = np.ascontiguousarray(r0, np.float64)
    r1 = np.ascontiguousarray(r1, np.float64)
    out = np.zeros((len(r0), 3), np.float32)
    w = _v3(omega_r)
    lib.orc_ggx_sample_n(float(roughness), _p(w), len(r0), _p(r0), _p(r1), _p(out))
    return out


def ggx_pdf_eval(roughness, omega_r, dirs, eta=1.5, lam=550.0):
    """MicrofacetGGX PDF (float64) and Eval (4 spectral values) of dirs in the tangent frame."""
    lib = _pin_sigs(load())
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    pdf = np.zeros(len(d), np.float64)
    ev = np.zeros((len(d), 4), np.float32)
    w = _v3(omega_r)
    lib.orc_ggx_pdf_eval_n(float(roughness), float(eta), float(lam), _p(w), len(d), _p(d), _p(pdf), _p(ev))
    return pdf, ev


def spectrum_rgb(rgb, lambdas):
    """ToRGB(FromRGB(rgb)) at each hero wavelength (float32 lambdas): (n, 3)."""
    lib = _pin_sigs(load())
    lam = np.ascontiguousarray(lambdas, np.float32).reshape(-1)
    out = np.zeros((len(lam), 3), np.float32)
    c = _v3(rgb)
    lib.orc_spectrum_rgb_n(_p(c), len(lam), _p(lam), _p(out))
    return out


def path_lobe_weight(kind, roughness, omega, r0, r1, eta=1.5, lam=550.0):
    """Path mode's continuation weight Eval/pdf of one lobe (kind 0 Oren-Nayar, 1 GGX) for sample pairs
    (r0, r1) in the tangent frame: (dirs (n, 3), weights (n, 4))."""
    lib = _pin_sigs(load())
    r0 = np.ascontiguousarray(r0, np.float64)
    r1 = np.ascontiguousarray(r1, np.float64)
    dirs = np.zeros((len(r0), 3), np.float32)
    out = np.zeros((len(r0), 4), np.float32)
    w = _v3(omega)
    lib.orc_path_lobe_weight_n(int(kind), float(roughness), float(eta), float(lam), _p(w), len(r0), _p(r0), _p(r1),
                               _p(dirs), _p(out))
    return dirs, out


def ggx_d_n(m, alpha):
    lib = _pin_sigs(load())
    mm = np.ascontiguousarray(m, np.float32).reshape(-1, 3)
    out = np.zeros(len(mm), np.float32)
    lib.orc_ggx_d_n(_p(mm), len(mm), float(alpha), _p(out))
    return out


def ggx_g1_n(w, m, alpha):
    lib = _pin_sigs(load())
    mm = np.ascontiguousarray(m, np.float32).reshape(-1, 3)
    out = np.zeros(len(mm), np.float32)
    ww = _v3(w)
    lib.orc_ggx_g1_n(_p(ww), _p(mm), len(mm), float(alpha), _p(out))
    return out


# ------------------------------------------------------------ guiding network (oracle_guide.cpp)

def _guide_sigs(lib):
    if getattr(lib, "_guide_ready", False):
        return lib
    P = C.c_void_p
    sig = {
        "ogd_nparams": (C.c_int, []),
        "ogd_forward": (None, [P, C.c_int64, P, P, P, C.c_int, P, P]),
        "ogd_mix_pdf": (None, [P, C.c_int64, P, P]),
        "ogd_nll_grad_raw": (None, [P, C.c_int64, P, P]),
        "ogd_nll": (None, [P, C.c_int64, P, P]),
        "ogd_loss_grad": (C.c_int, [P, C.c_int64, P, P, P, P, P, C.c_int, P, P]),
        "ogd_adam": (None, [P, P, P, P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "ogd_bf16": (C.c_float, [C.c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._guide_ready = True
    return lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


def guide_forward(params, pos, nrm, bounds, emulate=False):
    """FP64 forward pass: (raw (n,40), lobes (n,8,5))."""
    lib = _guide_sigs(load())
    pos, nrm = _f32(pos, (-1, 3)), _f32(nrm, (-1, 3))
    n = len(pos)
    raw = np.empty((n, 40), np.float32)
    lobes = np.empty((n, 8, 5), np.float32)
    lib.ogd_forward(_p(_f32(params)), n, _p(pos), _p(nrm), _p(_f32(bounds)), int(emulate), _p(raw), _p(lobes))
    return raw, lobes


def guide_mix_pdf(raw, dirs):
    lib = _guide_sigs(load())
    raw, dirs = _f32(raw, (-1, 40)), _f32(dirs, (-1, 3))
    out = np.empty(len(raw), np.float64)
    lib.ogd_mix_pdf(_p(raw), len(raw), _p(dirs), _p(out))
    return out


def guide_nll(raw, dirs):
    lib = _guide_sigs(load())
    raw = np.ascontiguousarray(raw, np.float64).reshape(-1, 40)
    dirs = _f32(dirs, (-1, 3))
    out = np.empty(len(raw), np.float64)
    lib.ogd_nll(_p(raw), len(raw), _p(dirs), _p(out))
    return out


def guide_nll_grad_raw(raw, dirs):
    lib = _guide_sigs(load())
    raw = np.ascontiguousarray(raw, np.float64).reshape(-1, 40)
    dirs = _f32(dirs, (-1, 3))
    out = np.empty_like(raw)
    lib.ogd_nll_grad_raw(_p(raw), len(raw), _p(dirs), _p(out))
    return out


def guide_loss_grad(params, pos, nrm, dirs, c, bounds, emulate=False):
    """Batch loss and mean gradient (float64, NPARAM); grad is None when sum c == 0."""
    lib = _guide_sigs(load())
    pos, nrm, dirs = _f32(pos, (-1, 3)), _f32(nrm, (-1, 3)), _f32(dirs, (-1, 3))
    c = _f32(c)
    loss = np.zeros(1, np.float64)
    grad = np.empty(lib.ogd_nparams(), np.float64)
    rc = lib.ogd_loss_grad(_p(_f32(params)), len(c), _p(pos), _p(nrm), _p(dirs), _p(c), _p(_f32(bounds)),
                           int(emulate), _p(loss), _p(grad))
    return float(loss[0]), (None if rc else grad)


def guide_adam(params, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays params, m, v."""
    lib = _guide_sigs(load())
    g = np.ascontiguousarray(grad, np.float64)
    lib.ogd_adam(_p(params), _p(m), _p(v), _p(g), len(params), int(t), lr, b1, b2, eps)


def bf16(x):
    lib = _guide_sigs(load())
    return np.array([lib.ogd_bf16(float(v)) for v in np.ravel(x)], np.float32).reshape(np.shape(x))


# ------------------------------------------------------------ path mode (oracle_pt.cpp)

def _pt_sigs(lib):
    if getattr(lib, "_pt_ready", False):
        return lib
    _render_sigs(lib)
    P = C.c_void_p
    sig = {
        "orc_pt_new": (P, [P, P, C.c_int32, P, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_int32, C.c_int32,
                           C.c_uint64]),
        "orc_pt_free": (None, [P]),
        "orc_pt_add_mesh_lights": (None, [P, P, P, C.c_int32, P, P, C.c_int32]),
        "orc_pt_pass": (C.c_int64, [P, C.c_uint32, P, P, C.c_int32]),
        "orc_pt_sincos2pi": (None, [C.c_float, P, P]),
        "orc_pt_set_fixed_guide": (None, [P, C.c_float, P]),
        "orc_pt_set_no_nee": (None, [P, C.c_int32]),
        "orc_pt_rnd": (C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
        "orc_pt_work": (None, [P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    lib._pt_ready = True
    return lib


class OraclePathTracer:
    """Plain path mode (alpha = 0) on the CPU.  desc: vermeer_amd.scenes.guided_room_scene() layout;
    cam: a vgpu_camera record (M, tan_theta_focal, aspect, focal)."""

    def __init__(self, desc, width, height, cam, max_depth=8, seed=1, leaf_size=16):
        self.lib = _pt_sigs(load())
        self.scene = OracleScene()
        self.w, self.h = width, height
        self._keep = []
        meshes = []
        for mesh in desc["meshes"]:
            if mesh.get("sphere"):
                raise ValueError("path mode takes triangle meshes only")
            v = np.ascontiguousarray(mesh["verts"], np.float32)
            ix = np.ascontiguousarray(mesh["idx"], np.uint32)
            sidx = mesh.get("shaderidx")
            sidx = None if sidx is None else np.ascontiguousarray(sidx, np.uint8)
            shaders = np.ascontiguousarray(mesh["shaders"], np.int32)
            self._keep += [v, ix, sidx, shaders]
            self.lib.orc_scene_add_shaded_mesh(self.scene.h, _p(v), len(v), _p(ix), len(ix), None, _p(sidx),
                                               _p(shaders), len(shaders))
            self.scene.meshes.append((v, ix))
            meshes.append((v, ix, sidx, shaders))
        self.scene.commit(leaf_size)
        mats = np.ascontiguousarray(desc["materials"]).view(np.float32).reshape(-1, 8)
        self._keep.append(mats)
        Mx = np.ascontiguousarray(cam["M"], np.float32)
        ttf, asp, foc = (float(np.asarray(cam[k]).reshape(-1)[0]) for k in ("tan_theta_focal", "aspect", "focal"))
        self.h_pt = self.lib.orc_pt_new(self.scene.h, _p(mats), len(mats), _p(Mx), ttf, asp, foc, width, height,
                                        max_depth, seed)
        for v, ix, sidx, shaders in meshes:
            self.lib.orc_pt_add_mesh_lights(self.h_pt, _p(v), _p(ix), len(ix), _p(sidx), _p(shaders), len(shaders))
        self.accum = np.zeros((height, width, 3), np.float64)
        self.passes = 0

    def __del__(self):
        if getattr(self, "h_pt", None):
            self.lib.orc_pt_free(self.h_pt)
            self.h_pt = None

    def set_fixed_guide(self, alpha, lobes):
        """Guided sampling with the same 8 lobes (mu xyz, kappa, weight) at every vertex."""
        self._lobes = np.ascontiguousarray(lobes, np.float32).reshape(40)
        self.lib.orc_pt_set_fixed_guide(self.h_pt, float(alpha), _p(self._lobes))

    def set_no_nee(self, on=True):
        """Diagnostic: sample directions only (no next-event estimation)."""
        self.lib.orc_pt_set_no_nee(self.h_pt, int(bool(on)))

    def traversal_work(self):
        """Traversal work since the last call (then reset): {"closest": (rays, node visits, triangle
        tests), "shadow": (...)}."""
        w = np.zeros(6, np.int64)
        self.lib.orc_pt_work(self.h_pt, _p(w))
        return {"closest": tuple(int(x) for x in w[:3]), "shadow": tuple(int(x) for x in w[3:])}

    def render_pass(self, pass_index, nthreads=0, want_sample=False):
        samp = np.zeros((self.h, self.w, 3), np.float32) if want_sample else None
        n = self.lib.orc_pt_pass(self.h_pt, pass_index, _p(self.accum), _p(samp), nthreads)
        self.passes += 1
        return n, samp


def pt_sincos2pi(u):
    lib = _pt_sigs(load())
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib.orc_pt_sincos2pi(float(u), _p(s), _p(c))
    return float(s[0]), float(c[0])


def pt_rnd(pixel, pass_index, dim, seed):
    return _pt_sigs(load()).orc_pt_rnd(pixel, pass_index, dim, seed)


# ------------------------------------------------------------ output drivers (oracle_output.cpp)

def _output_sigs(lib):
    P = C.c_void_p
    for name, res, args in (("orc_film_to_rgbe", None, [P, C.c_int32, C.c_int32, P]),
                            ("orc_write_hdr", C.c_int, [P, C.c_int32, C.c_int32, C.c_char_p]),
                            ("orc_write_float", C.c_int, [P, C.c_int32, C.c_int32, C.c_char_p])):
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def film_to_rgbe(film):
    """convertRGBToRGBE over a (h, w, 3) float32 film, in the file's scanline order: (h, w, 4) uint8."""
    lib = _output_sigs(load())
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    out = np.zeros((h, w, 4), np.uint8)
    lib.orc_film_to_rgbe(_p(f), w, h, _p(out))
    return out


def write_hdr(film, path):
    lib = _output_sigs(load())
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    assert lib.orc_write_hdr(_p(f), w, h, str(path).encode()) == 0


def write_float(film, path):
    lib = _output_sigs(load())
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    assert lib.orc_write_float(_p(f), w, h, str(path).encode()) == 0
