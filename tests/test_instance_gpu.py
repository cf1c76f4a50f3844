"""GPU parity of instances, procs and mesh transforms against the oracle, through the C ABI.

Workload: the instance scene (vermeer_amd.scenes.instance_scene): 12 GeomInstances of three hidden
prototype meshes under random rotation, non-uniform scale and translation (every third with two keys:
moving and turning over the shutter), a Proc over two hidden meshes under a rotation, and a plank with
its own two-key Transform, in the feature scene's room with its disk light.  Rays: 200k random rays in
the room at random times, closest hit and shadow; renders in the reference's shading and in path mode.
Bar: hit prim, geom slot, t, u and v bit-identical; every per-pass sample and the film bit-identical
(both sides on the IEEE 1/sqrt estimate in Vec3Normalize).
"""
import numpy as np
import pytest

import oracle
import vermeer_amd as va
from vermeer_amd import render, scenes

pytestmark = pytest.mark.gpu


def _camera(desc, w, h):
    c = desc["camera"]
    M, ttf = oracle.camera_lookat(c["from"], c["to"], c["up"], c["fov"], c["focal"])
    cam = np.zeros((), va._lib.CAMERA_DTYPE)
    cam["M"] = M
    cam["tan_theta_focal"] = ttf
    cam["aspect"] = w / h
    cam["focal"] = c["focal"]
    return cam, M, ttf


@pytest.fixture(scope="module")
def inst():
    print("\n[case] instance scene", flush=True)
    return scenes.instance_scene()


def _room_rays(n=200_000, seed=1):
    g = np.random.default_rng(seed)
    r = np.zeros(n, va.RAY_DTYPE)
    r["P"] = g.uniform((-0.9, 0.1, -0.9), (0.9, 1.9, 0.9), (n, 3)).astype(np.float32)
    d = g.normal(size=(n, 3))
    r["D"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmax"] = np.inf
    r["time"] = g.uniform(0, 1, n).astype(np.float32)
    return r


def test_instance_scene_traversal_bitexact(inst):
    cam, M, ttf = _camera(inst, 16, 16)
    g = render.Renderer(inst, 16, 16, cam)
    o = oracle.OracleRenderer(inst, 16, 16, M, ttf, inst["camera"]["focal"])
    rays = _room_rays()
    for shadow in (False, True):
        print(f"[case] instance scene trace shadow={shadow}", flush=True)
        hits, geoms, _ = g.scene.trace(rays, shadow=shadow)
        ref, _ = o.scene.trace(rays, shadow=shadow)
        assert np.array_equal(hits["prim"], ref["prim"])
        assert np.array_equal(geoms, ref["geom"])
        assert hits["t"].tobytes() == ref["t"].tobytes()
        assert hits["u"].tobytes() == ref["u"].tobytes() and hits["v"].tobytes() == ref["v"].tobytes()
        if not shadow:
            closest = geoms
    # every geom of the top level (box, disk, plank, the instances, the proc) holds some closest hit
    assert len(np.unique(closest[closest >= 0])) == len(inst["meshes"]) - 5 + len(inst["instances"]) + 1


def test_instance_scene_rejects_counters_and_bad_transforms(inst):
    cam, _, _ = _camera(inst, 16, 16)
    g = render.Renderer(inst, 16, 16, cam)
    with pytest.raises(va.VgpuError):
        g.scene.trace(_room_rays(64), want_counters=True)
    sc = va.Scene(0)
    v, f = scenes.uv_sphere(1.0, 8, 4)
    m = sc.add_geom(va.PolyMesh(v, f))
    sph = sc.add_sphere((0, 0, 0), 1.0, 0)
    with pytest.raises(va.VgpuError):
        sc.set_transform(sph, scenes.trs_matrix())               # an analytic sphere has no Transform
    with pytest.raises(va.VgpuError):
        sc.set_transform(0, np.zeros(16, np.float32))            # singular
    i = sc.add_instance(0, scenes.trs_matrix((1, 0, 0)), (-1, -1, -1), (2, 1, 1))
    with pytest.raises(va.VgpuError):
        sc.add_instance(i, scenes.trs_matrix(), (-1, -1, -1), (1, 1, 1))   # instances hold meshes and spheres
    with pytest.raises(va.VgpuError):
        sc.add_proc([], None, (-1, -1, -1), (1, 1, 1))
    del m


@pytest.mark.parametrize("depth", [0, 4])
def test_instance_scene_render_bitexact(inst, depth):
    oracle.set_normalize_ieee(True)
    try:
        w = h = 128
        cam, M, ttf = _camera(inst, w, h)
        g = render.Renderer(inst, w, h, cam)
        o = oracle.OracleRenderer(inst, w, h, M, ttf, inst["camera"]["focal"])
        g.set_path_depth(depth)
        o.set_path_depth(depth)
        scr = scenes.pixel_scrambles(w, h, seed=19)
        for it in range(1, 4):
            print(f"[case] instance scene render depth {depth} pass {it}", flush=True)
            ng, gs = g.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            bad = ~((gs.view(np.uint32) == os_.view(np.uint32)) | (np.isnan(gs) & np.isnan(os_)))
            assert not bad.any(), f"pass {it}: {int(bad.any(axis=2).sum())} pixels differ, first {np.argwhere(bad)[:3]}"
            assert ng == no
        assert g.film.mean() > 0.02
    finally:
        oracle.set_normalize_ieee(False)
