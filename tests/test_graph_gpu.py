"""HIP-graph passes and the autotuner, through the C ABI.

- VGPU_GRAPH: a render pass and a path-mode pass captured into a HIP graph and replayed (updated in
  place from pass to pass) give exactly what eager launches give, pass after pass, with eager work of
  the program between the replays and switches between graph and eager passes on one object.
- vgpu_scene_autotune: the configuration it picks (block, LDS depth, waves, sort, refill threshold)
  traces bit-exactly as the oracle does, and every refill threshold gives the same hits.
"""
import numpy as np
import pytest
import torch

import oracle
import vermeer_amd as va
from vermeer_amd import pathtrace, render, scenes

pytestmark = pytest.mark.gpu


def _scr(W, seed):
    return torch.from_numpy(scenes.pixel_scrambles(W, W, seed=seed).view(np.int64).reshape(-1).copy()).cuda()


@pytest.mark.parametrize("scene", ["box_spheres", "feature"])
def test_render_pass_graph_equals_eager(scene):
    desc = scenes.box_spheres_scene(nseg=32, nring=16) if scene == "box_spheres" else scenes.feature_scene(nseg=24, nring=12)
    c = desc["camera"]
    W = 128
    cam = render.lookat_camera(c["from"], c["to"], c["up"], c["fov"], c["focal"], W, W)
    print(f"\n[case] render graph vs eager, {scene} {W}x{W}", flush=True)
    Re = render.Renderer(desc, W, W, cam)
    Rg = render.Renderer(desc, W, W, cam)
    scr = _scr(W, 4)
    st = torch.cuda.current_stream().cuda_stream
    fe = torch.zeros(W * W * 3, dtype=torch.float32, device="cuda")
    fg = torch.zeros_like(fe)
    se = torch.zeros_like(fe)
    sg = torch.zeros_like(fe)
    for it in range(1, 9):
        Re.render_pass_device(scr.data_ptr(), it, fe.data_ptr(), d_sample=se.data_ptr(), stream=st)
        Rg.render_pass_device(scr.data_ptr(), it, fg.data_ptr(), d_sample=sg.data_ptr(), stream=st,
                              graph=(it != 5))          # pass 5 eager on the graph renderer
        junk = torch.randn(4096, device="cuda").sum()   # the program's own eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        assert torch.equal(fe.view(torch.int32), fg.view(torch.int32)), f"pass {it}: film differs"
        assert torch.equal(se.view(torch.int32), sg.view(torch.int32)), f"pass {it}: samples differ"
    Re.scene.status(st)
    Rg.scene.status(st)
    assert float(fe.mean()) > 0
    Re.close()
    Rg.close()


def test_path_pass_graph_equals_eager():
    d = scenes.guided_room_scene(res=5, nboxes=2000)
    W = 96
    print(f"\n[case] path mode graph vs eager {W}x{W}", flush=True)
    pe = pathtrace.PathTracer(d, W, W, max_depth=8, seed=3)
    pg = pathtrace.PathTracer(d, W, W, max_depth=8, seed=3)
    st = torch.cuda.current_stream().cuda_stream
    ae = torch.zeros(W * W * 3, dtype=torch.float64, device="cuda")
    ag = torch.zeros_like(ae)
    for p in range(8):
        pe.render_pass_device(ae.data_ptr(), p, stream=st)
        pg.render_pass_device(ag.data_ptr(), p, stream=st, graph=(p % 3 != 2))
        junk = torch.ones(1000, device="cuda").cumsum(0)[-1]
        torch.cuda.synchronize()
        assert float(junk) == 1000.0
        assert torch.equal(ae.view(torch.int64), ag.view(torch.int64)), f"pass {p}: accumulation differs"
    assert pe.stats()[0] == pg.stats()[0]
    assert float(ae.max()) > 0
    pe.close()
    pg.close()


def test_graph_flag_needs_device_pointers():
    desc = scenes.box_spheres_scene(nseg=8, nring=4)
    c = desc["camera"]
    cam = render.lookat_camera(c["from"], c["to"], c["up"], c["fov"], c["focal"], 16, 16)
    R = render.Renderer(desc, 16, 16, cam)
    lib = va._lib.load()
    film = np.zeros(16 * 16 * 3, np.float32)
    scr = scenes.pixel_scrambles(16, 16)
    rc = lib.vgpu_render_pass(R._h, va._lib.VGPU_GRAPH, va._lib.ptr(scr), 1, va._lib.ptr(film), None, None, None)
    assert rc == va._lib.VGPU_ERR_ARG
    R.close()


def test_autotune_choice_traces_like_the_oracle():
    verts, idx = scenes.icosphere_field(count=256)[:2]
    g = va.Scene(0)
    g.add_geom(va.PolyMesh(verts, idx))
    g.pre_render()
    o = oracle.OracleScene()
    o.add_mesh(verts, idx)
    o.commit()
    rays = scenes.bounce_rays(verts, idx, 262_144, seed=5)[0]
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    d_hits = torch.empty(len(rays) * 16, dtype=torch.uint8, device="cuda")
    print("\n[case] autotune on 262k bounce rays over 327k triangles", flush=True)
    cfg, ms = g.autotune(d_rays.data_ptr(), len(rays), d_hits.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
    print(f"[case] tuned {cfg} {ms:.3f} ms", flush=True)
    assert cfg == g.trace_config() and ms > 0
    assert cfg["refill"] in (16, 32, 40, 48, 56, 64)
    ref, rcnt = o.trace(rays, counters=True)
    hits, geoms, cnt = g.trace(rays, want_counters=True)
    assert np.array_equal(hits["prim"], ref["prim"]) and hits["t"].tobytes() == ref["t"].tobytes()
    assert hits["u"].tobytes() == ref["u"].tobytes() and hits["v"].tobytes() == ref["v"].tobytes()
    assert np.array_equal(cnt, rcnt)
    for refill in (1, 16, 64):
        g.set_trace_config(cfg["block"], cfg["lds_stack"], cfg["waves_per_cu"], cfg["sort"], refill)
        h2, _, c2 = g.trace(rays, want_counters=True)
        assert h2.tobytes() == hits.tobytes() and np.array_equal(c2, cnt), refill
    g.close()
