"""GPU parity of the renderer (reference semantics) against the oracle, through the C ABI.

Workloads: the box-and-spheres scene (5-wall box, 8 noise-displaced 64x32 UV spheres, 0.5 x 0.5 light
as two triangle lights), 512 x 512, 16 passes of deterministic QMC (RasterXY / VdC / Sobol with a
seeded scramble array); and the feature scene (conductor and perfect-mirror spheres, a Disk light and a
Sphere light with its analytic sphere) under the box, Gaussian and Airy pixel filters.

Bar: with both sides using the IEEE estimate in Vec3Normalize, every per-pass sample and the film are
bit-identical (stricter than the 1e-4 relative bar).  With the oracle on the reference's RSQRTSS
estimate (Go on this CPU), pixels still agree within 1e-4 relative except where a 1-ulp change of a
direction flips a discrete outcome; that fraction is bounded.
"""
import numpy as np
import pytest

import oracle
import vermeer_amd as va
from vermeer_amd import render, scenes

pytestmark = pytest.mark.gpu

W = H = 512
PASSES = 16


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
def box_scene():
    print("\n[case] box_spheres scene", flush=True)
    return scenes.box_spheres_scene()


def _nan_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def test_render_box_spheres_bitexact(box_scene):
    oracle.set_normalize_ieee(True)
    try:
        cam, M, ttf = _camera(box_scene, W, H)
        g = render.Renderer(box_scene, W, H, cam)
        o = oracle.OracleRenderer(box_scene, W, H, M, ttf, box_scene["camera"]["focal"])
        scr = scenes.pixel_scrambles(W, H)
        for it in range(1, PASSES + 1):
            print(f"[case] render pass {it}", flush=True)
            ng, gs = g.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        assert _nan_equal(g.film, o.film)
        assert g.film.mean() > 0.05
    finally:
        oracle.set_normalize_ieee(False)


def test_render_within_1e4_of_reference_normalize(box_scene):
    """Oracle on RSQRTSS (the reference's arithmetic on this host CPU): the film agrees within 1e-4
    relative (absolute floor 1e-3) on all but a small fraction of pixels: 1.84% measured on the MI355X
    hosts, where a 1-ulp direction change flips a shadow ray grazing an edge (the IEEE-mode test above
    is bit-exact)."""
    cam, M, ttf = _camera(box_scene, 256, 256)
    g = render.Renderer(box_scene, 256, 256, cam)
    o = oracle.OracleRenderer(box_scene, 256, 256, M, ttf, box_scene["camera"]["focal"])
    scr = scenes.pixel_scrambles(256, 256, seed=5)
    for it in range(1, 9):
        g.render_pass(scr, it)
        o.render_pass(scr, it)
    a, b = g.film.astype(np.float64), o.film.astype(np.float64)
    ok = np.isnan(a) & np.isnan(b)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    frac_bad = float(np.mean(~(ok | (rel <= 1e-4))))
    print(f"[case] RSQRTSS oracle: fraction of film values beyond 1e-4: {frac_bad:.4f}", flush=True)
    assert frac_bad < 0.025


def test_render_device_pointer_path(box_scene):
    import torch
    oracle.set_normalize_ieee(True)
    try:
        w = h = 128
        cam, M, ttf = _camera(box_scene, w, h)
        g = render.Renderer(box_scene, w, h, cam)
        o = oracle.OracleRenderer(box_scene, w, h, M, ttf, box_scene["camera"]["focal"])
        scr = scenes.pixel_scrambles(w, h, seed=7)
        d_scr = torch.from_numpy(scr.view(np.int64).reshape(-1).copy()).cuda()
        d_film = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        for it in (1, 2, 3):
            g.render_pass_device(d_scr.data_ptr(), it, d_film.data_ptr(), stream=st)
            o.render_pass(scr, it)
        g.scene.status(st)
        assert _nan_equal(d_film.cpu().numpy().reshape(h, w, 3), o.film)
    finally:
        oracle.set_normalize_ieee(False)


def test_render_rejects_invalid_materials_and_lights(box_scene):
    mats = box_scene["materials"].copy()
    mats[1]["fresnel_model"] = 3               # neither Dielectric nor Metal
    d = dict(box_scene, materials=mats)
    cam, _, _ = _camera(d, 64, 64)
    with pytest.raises(va.VgpuError) as ei:
        render.Renderer(d, 64, 64, cam)
    assert ei.value.code == va._lib.VGPU_ERR_ARG
    lights = [dict(box_scene["lights"][0], type="sphere", P=(0.0, 1.0, 0.0), radius=0.1)]   # geom is a mesh
    d = dict(box_scene, lights=lights)
    with pytest.raises(va.VgpuError) as ei:
        render.Renderer(d, 64, 64, cam)
    assert ei.value.code == va._lib.VGPU_ERR_ARG


@pytest.fixture(scope="module")
def feature_scene():
    print("\n[case] feature scene (metal / mirror spheres, disk and sphere lights)", flush=True)
    return scenes.feature_scene()


@pytest.mark.parametrize("filt", ["box", "gaussian", "airy"])
def test_render_feature_scene_bitexact(feature_scene, filt):
    """Conductor Fresnel, perfect mirrors to level 3, Disk and Sphere lights, an analytic sphere geom and
    the three pixel filters, against the oracle: every per-pass sample bit-identical."""
    oracle.set_normalize_ieee(True)
    try:
        w = h = 160
        cam, M, ttf = _camera(feature_scene, w, h)
        g = render.Renderer(feature_scene, w, h, cam)
        o = oracle.OracleRenderer(feature_scene, w, h, M, ttf, feature_scene["camera"]["focal"])
        g.set_filter(filt)
        o.set_filter(filt)
        scr = scenes.pixel_scrambles(w, h, seed=21)
        for it in range(1, 5):
            print(f"[case] feature scene filter {filt} pass {it}", flush=True)
            ng, gs = g.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            bad = ~((gs.view(np.uint32) == os_.view(np.uint32)) | (np.isnan(gs) & np.isnan(os_)))
            assert not bad.any(), f"pass {it}: {int(bad.any(axis=2).sum())} pixels differ, first {np.argwhere(bad)[:3]}"
        assert _nan_equal(g.film, o.film)
        assert g.film.mean() > 0.02
    finally:
        oracle.set_normalize_ieee(False)


def test_render_thin_lens_bitexact(feature_scene):
    oracle.set_normalize_ieee(True)
    try:
        w = h = 128
        cam, M, ttf = _camera(feature_scene, w, h)
        cam["radius"] = 0.06
        g = render.Renderer(feature_scene, w, h, cam)
        o = oracle.OracleRenderer(feature_scene, w, h, M, ttf, feature_scene["camera"]["focal"])
        o.set_lens(0.06)
        scr = scenes.pixel_scrambles(w, h, seed=13)
        for it in range(1, 4):
            print(f"[case] thin lens pass {it}", flush=True)
            _, gs = g.render_pass(scr, it, want_sample=True)
            _, os_ = o.render_pass(scr, it, want_sample=True)
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
    finally:
        oracle.set_normalize_ieee(False)


def test_render_feature_scene_device_pointer_path(feature_scene):
    import torch
    oracle.set_normalize_ieee(True)
    try:
        w = h = 96
        cam, M, ttf = _camera(feature_scene, w, h)
        g = render.Renderer(feature_scene, w, h, cam)
        o = oracle.OracleRenderer(feature_scene, w, h, M, ttf, feature_scene["camera"]["focal"])
        g.set_filter("gaussian")
        o.set_filter("gaussian")
        scr = scenes.pixel_scrambles(w, h, seed=8)
        d_scr = torch.from_numpy(scr.view(np.int64).reshape(-1).copy()).cuda()
        d_film = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        for it in (1, 2, 3):
            g.render_pass_device(d_scr.data_ptr(), it, d_film.data_ptr(), stream=st)
            o.render_pass(scr, it)
        g.scene.status(st)
        assert _nan_equal(d_film.cpu().numpy().reshape(h, w, 3), o.film)
    finally:
        oracle.set_normalize_ieee(False)


@pytest.fixture(scope="module")
def moving_scene():
    print("\n[case] moving scene", flush=True)
    return scenes.moving_scene(nseg=32, nring=16)


def test_render_moving_scene_bitexact(moving_scene):
    """Moving meshes (per-key normals; one with the reference's slot indexing and single-key normals)
    traced and shaded at each sample's time, perfect-mirror recursion and shadow rays carrying it."""
    oracle.set_normalize_ieee(True)
    try:
        w = h = 256
        cam, M, ttf = _camera(moving_scene, w, h)
        g = render.Renderer(moving_scene, w, h, cam)
        o = oracle.OracleRenderer(moving_scene, w, h, M, ttf, moving_scene["camera"]["focal"])
        scr = scenes.pixel_scrambles(w, h, seed=21)
        for it in range(1, 9):
            print(f"[case] moving scene pass {it}", flush=True)
            ng, gs = g.render_pass(scr, it, want_sample=True)
            no, os_ = o.render_pass(scr, it, want_sample=True)
            assert ng == no, f"pass {it}: shadow rays {ng} vs {no}"
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        assert _nan_equal(g.film, o.film)
        assert g.film.mean() > 0.02
    finally:
        oracle.set_normalize_ieee(False)


def test_render_moving_scene_lens_filter_device_path(moving_scene):
    import torch
    oracle.set_normalize_ieee(True)
    try:
        w = h = 96
        cam, M, ttf = _camera(moving_scene, w, h)
        cam["radius"] = 0.04
        g = render.Renderer(moving_scene, w, h, cam)
        o = oracle.OracleRenderer(moving_scene, w, h, M, ttf, moving_scene["camera"]["focal"])
        o.set_lens(0.04)
        g.set_filter("gaussian")
        o.set_filter("gaussian")
        scr = scenes.pixel_scrambles(w, h, seed=9)
        d_scr = torch.from_numpy(scr.view(np.int64).reshape(-1).copy()).cuda()
        d_film = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        for it in (1, 2, 3):
            print(f"[case] moving scene device pass {it}", flush=True)
            g.render_pass_device(d_scr.data_ptr(), it, d_film.data_ptr(), stream=st)
            o.render_pass(scr, it)
        g.scene.status(st)
        assert _nan_equal(d_film.cpu().numpy().reshape(h, w, 3), o.film)
    finally:
        oracle.set_normalize_ieee(False)


def test_add_shaded_motion_mesh_rejects_bad_arguments():
    import ctypes as C
    from vermeer_amd.accel import Scene
    lib = va._lib.load()
    s = Scene(0)
    v = np.zeros((2, 3, 3), np.float32)
    v[:, 1, 0] = 1.0
    v[:, 2, 1] = 1.0
    ix = np.array([[0, 1, 2]], np.uint32)
    sh = np.array([0], np.int32)
    g = C.c_int32(-1)
    p = va._lib.ptr
    assert lib.vgpu_scene_add_shaded_motion_mesh(s._h, p(v), 3, 1, p(ix), 1, None, 0, None, p(sh), 1, 0, C.byref(g)) != 0
    assert lib.vgpu_scene_add_shaded_motion_mesh(s._h, p(v), 3, 2, p(ix), 1, None, 0, None, p(sh), 1, 4, C.byref(g)) != 0
    assert lib.vgpu_scene_add_shaded_motion_mesh(s._h, p(v), 3, 2, p(ix), 1, None, 0, None, None, 0, 0, C.byref(g)) != 0
    bad_sidx = np.array([3], np.uint8)
    assert lib.vgpu_scene_add_shaded_motion_mesh(s._h, p(v), 3, 2, p(ix), 1, None, 0, p(bad_sidx), p(sh), 1, 0,
                                                 C.byref(g)) != 0
    assert lib.vgpu_scene_add_shaded_motion_mesh(s._h, p(v), 3, 2, p(ix), 1, None, 0, None, p(sh), 1, 0, C.byref(g)) == 0
    assert g.value == 0
    s.close()


def test_render_camera_motion_keys_bitexact(moving_scene):
    """Keyed LookAt camera (three From keys, two Roll keys) over the moving scene: each sample's camera
    matrix is recomposed from the lerped decomposition at its time, on the GPU as in the oracle."""
    oracle.set_normalize_ieee(True)
    try:
        w = h = 128
        c = moving_scene["camera"]
        cam, M, ttf = _camera(moving_scene, w, h)
        keys = oracle.camera_lookat_keys([(0.0, 1.0, 3.4), (0.4, 1.1, 3.3), (0.9, 1.25, 3.1)], [c["to"]], c["up"],
                                         roll=[0.0, 0.15])
        g = render.Renderer(moving_scene, w, h, cam)
        o = oracle.OracleRenderer(moving_scene, w, h, M, ttf, c["focal"])
        g.set_camera_keys(keys)
        o.set_camera_keys(keys)
        scr = scenes.pixel_scrambles(w, h, seed=31)
        for it in range(1, 5):
            print(f"[case] camera keys pass {it}", flush=True)
            _, gs = g.render_pass(scr, it, want_sample=True)
            _, os_ = o.render_pass(scr, it, want_sample=True)
            assert _nan_equal(gs, os_), f"pass {it}: samples differ"
        # one key: the library's host decomposition sets the static matrix
        g1 = render.Renderer(moving_scene, w, h, cam)
        o1 = oracle.OracleRenderer(moving_scene, w, h, M, ttf, c["focal"])
        g1.set_camera_keys(keys[2:])
        o1.set_camera_keys(keys[2:])
        _, gs = g1.render_pass(scr, 1, want_sample=True)
        _, os_ = o1.render_pass(scr, 1, want_sample=True)
        assert _nan_equal(gs, os_)
        with pytest.raises(Exception):
            g1.set_camera_keys(np.full((1, 16), np.nan, np.float32))
    finally:
        oracle.set_normalize_ieee(False)


def test_camera_keys_reset_and_limits(feature_scene):
    """Zero keys return to the creation-time camera matrix (also after a one-key call replaced it);
    more than 64 keys are refused."""
    oracle.set_normalize_ieee(True)
    try:
        w = h = 64
        c = feature_scene["camera"]
        cam, M, ttf = _camera(feature_scene, w, h)
        keys = oracle.camera_lookat_keys([(0.3, 1.2, 3.2)], [c["to"]], c["up"])
        g = render.Renderer(feature_scene, w, h, cam)
        o = oracle.OracleRenderer(feature_scene, w, h, M, ttf, c["focal"])
        scr = scenes.pixel_scrambles(w, h, seed=41)
        g.set_camera_keys(keys)
        g.set_camera_keys(np.zeros((0, 16), np.float32))
        _, gs = g.render_pass(scr, 1, want_sample=True)
        _, os_ = o.render_pass(scr, 1, want_sample=True)
        assert _nan_equal(gs, os_)
        with pytest.raises(Exception):
            g.set_camera_keys(np.repeat(keys, 65, axis=0))
    finally:
        oracle.set_normalize_ieee(False)
