"""GPU parity of the renderer's path mode (vgpu_render_set_path_depth) against the oracle, through the C ABI.

Path mode shades every vertex of a path with ShaderStd as the reference shades its camera hit
(EvaluateLightSamples per light and lobe, Tri / Disk / Sphere lights, Oren-Nayar, GGX and mirror lobes,
spectral colour), continues each path with one ray sampled from a lobe picked by its weight, and draws
every dimension from the reference's ldseq sequences (RasterXY, VdC, Sobol with per-level scrambles).

Workloads: the box-and-spheres scene at its full render size (512 x 512, 16 passes, max depth 5); the
feature scene (conductor and mirror spheres, Disk and Sphere lights) under the Gaussian filter at depth 5;
the moving scene at depth 3; the emissive furnace cube with no lights; a pass replayed as a HIP graph.
Bar: with both sides on the IEEE 1/sqrt estimate in Vec3Normalize, every per-pass sample and the film are
bit-identical.
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


def _pair(desc, w, h, depth, filt="box"):
    cam, M, ttf = _camera(desc, w, h)
    g = render.Renderer(desc, w, h, cam)
    o = oracle.OracleRenderer(desc, w, h, M, ttf, desc["camera"]["focal"])
    g.set_path_depth(depth)
    o.set_path_depth(depth)
    if filt != "box":
        g.set_filter(filt)
        o.set_filter(filt)
    return g, o


def _diff(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return ~((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b)))


def _run(g, o, scr, passes, label):
    nshadow = 0
    for it in range(1, passes + 1):
        print(f"[case] {label} pass {it}", flush=True)
        ng, gs = g.render_pass(scr, it, want_sample=True)
        no, os_ = o.render_pass(scr, it, want_sample=True)
        bad = _diff(gs, os_)
        assert not bad.any(), f"{label} pass {it}: {int(bad.any(axis=2).sum())} pixels differ, first {np.argwhere(bad)[:3]}"
        assert ng == no, f"{label} pass {it}: shadow rays {ng} vs {no}"
        nshadow += no
    assert not _diff(g.film, o.film).any()
    # the device ray counters (vgpu_render_stats) agree with the oracle's shadow rays, and every pass
    # traced at least one camera ray per pixel
    closest, shadow = g.stats()
    assert shadow == nshadow, f"{label}: device shadow rays {shadow} vs {nshadow}"
    assert closest >= passes * g.width * g.height


def test_path_box_spheres_full_size_depth5_bitexact():
    """The box-and-spheres render at 512 x 512, 16 passes, max depth 5 (GGX roughness 0.05-0.5, Lambert
    albedo 0.2-0.8, the 0.5 x 0.5 ceiling light as two triangle lights)."""
    oracle.set_normalize_ieee(True)
    try:
        d = scenes.box_spheres_scene()
        g, o = _pair(d, 512, 512, 5)
        scr = scenes.pixel_scrambles(512, 512)
        _run(g, o, scr, 16, "box_spheres depth 5")
        # the bounces add light: brighter than the direct-only reference render of the same passes
        ref = oracle.OracleRenderer(d, 512, 512, *_camera(d, 512, 512)[1:], d["camera"]["focal"])
        for it in range(1, 5):
            ref.render_pass(scr, it)
        assert g.film.mean() > 1.3 * ref.film.mean()
    finally:
        oracle.set_normalize_ieee(False)


def test_path_feature_scene_gaussian_depth5_bitexact():
    """Conductor and mirror spheres (mirror continuations), Disk and Sphere lights with their geometry
    (emission dropped after smooth bounces), the Gaussian filter."""
    oracle.set_normalize_ieee(True)
    try:
        d = scenes.feature_scene()
        g, o = _pair(d, 160, 160, 5, filt="gaussian")
        _run(g, o, scenes.pixel_scrambles(160, 160, seed=17), 4, "feature depth 5")
        assert g.film.mean() > 0.02
    finally:
        oracle.set_normalize_ieee(False)


def test_path_moving_scene_depth3_bitexact():
    """Moving meshes: every continuation carries its sample's time."""
    oracle.set_normalize_ieee(True)
    try:
        d = scenes.moving_scene(nseg=32, nring=16)
        g, o = _pair(d, 128, 128, 3)
        _run(g, o, scenes.pixel_scrambles(128, 128, seed=23), 3, "moving depth 3")
    finally:
        oracle.set_normalize_ieee(False)


def test_path_furnace_cube_without_lights_bitexact():
    import test_render_path_oracle as pins
    oracle.set_normalize_ieee(True)
    try:
        d = pins._furnace_cube(np.array([0.5, 0.6, 0.7], np.float32), np.array([1.0, 0.8, 0.6], np.float32))
        g, o = _pair(d, 64, 64, 5)
        _run(g, o, scenes.pixel_scrambles(64, 64, seed=8), 3, "furnace depth 5")
        assert g.film.mean() > 1.0
    finally:
        oracle.set_normalize_ieee(False)


def test_path_depth_switches_on_one_renderer():
    """Depth 0 (the reference), 5 and 1 on one renderer: depth 0 and depth 1 give the reference's film on
    a scene without mirrors; depth 5 the oracle's path-mode film."""
    oracle.set_normalize_ieee(True)
    try:
        d = scenes.box_spheres_scene()
        w = h = 96
        cam, M, ttf = _camera(d, w, h)
        g = render.Renderer(d, w, h, cam)
        scr = scenes.pixel_scrambles(w, h, seed=31)
        films = {}
        for depth in (0, 5, 1):
            o = oracle.OracleRenderer(d, w, h, M, ttf, d["camera"]["focal"])
            o.set_path_depth(depth)
            g.set_path_depth(depth)
            g.film[:] = 0
            for it in (1, 2):
                g.render_pass(scr, it)
                o.render_pass(scr, it)
            assert not _diff(g.film, o.film).any(), f"depth {depth}"
            films[depth] = g.film.copy()
        assert not _diff(films[0], films[1]).any()
        # depth 0 and 1 trace exactly one closest-hit ray per pixel and pass (2 + 2 passes), depth 5 more
        closest, _ = g.stats()
        assert closest > 6 * w * h
        with pytest.raises(va.VgpuError):
            g.set_path_depth(17)
    finally:
        oracle.set_normalize_ieee(False)


def test_path_pass_as_graph_equals_eager():
    import torch
    oracle.set_normalize_ieee(True)
    try:
        d = scenes.box_spheres_scene()
        w = h = 128
        cam, M, ttf = _camera(d, w, h)
        g = render.Renderer(d, w, h, cam)
        g.set_path_depth(5)
        o = oracle.OracleRenderer(d, w, h, M, ttf, d["camera"]["focal"])
        o.set_path_depth(5)
        scr = scenes.pixel_scrambles(w, h, seed=5)
        d_scr = torch.from_numpy(scr.view(np.int64).reshape(-1).copy()).cuda()
        d_film = torch.zeros(w * h * 3, dtype=torch.float32, device="cuda")
        st = torch.cuda.current_stream().cuda_stream
        for it in range(1, 6):
            print(f"[case] path graph pass {it}", flush=True)
            g.render_pass_device(d_scr.data_ptr(), it, d_film.data_ptr(), stream=st, graph=(it != 3))
            torch.cuda.synchronize()
            o.render_pass(scr, it)
            assert not _diff(d_film.cpu().numpy().reshape(h, w, 3), o.film).any(), f"pass {it}"
        g.scene.status(st)
    finally:
        oracle.set_normalize_ieee(False)
