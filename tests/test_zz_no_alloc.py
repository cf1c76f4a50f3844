"""GPU parity of the allocation-free traversal entry point (vgpu_trace_noalloc).

It must give exactly what vgpu_trace gives on the same device buffers: hits,
geom slots and per-ray traversal counters, bit for bit, for closest and any
hit, on two calls in a row (the second smaller, so the scratch array holds the
first call's spill), on a static and a motion scene, and with a launch shape
whose spill columns do not all fit the scratch array.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import vermeer_amd as va
from vermeer_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _run(g, fn_noalloc, d_rays, n, shadow, stream):
    """Both entry points on the same rays; returns (reference, no-alloc) tuples of host arrays."""
    lib = _lib.load()
    out = []
    for noalloc in (False, True):
        d_hits = torch.full((n * 16,), 0x5A, dtype=torch.uint8, device="cuda")
        d_geoms = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        d_cnt = torch.full((n * 4,), -7, dtype=torch.int32, device="cuda")
        d_status = torch.full((1,), 99, dtype=torch.int32, device="cuda")
        flags = (_lib.VGPU_TRACE_SHADOW if shadow else 0) | _lib.VGPU_TRACE_DEVICE_PTRS
        if noalloc:
            _lib.check(lib.vgpu_trace_noalloc(g._h, flags, C.c_void_p(d_rays.data_ptr()), n,
                                              C.c_void_p(d_hits.data_ptr()), C.c_void_p(d_geoms.data_ptr()),
                                              C.c_void_p(d_cnt.data_ptr()), C.c_void_p(d_status.data_ptr()),
                                              C.c_void_p(stream)))
        else:
            _lib.check(lib.vgpu_trace(g._h, flags, C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(d_hits.data_ptr()),
                                      C.c_void_p(d_geoms.data_ptr()), C.c_void_p(d_cnt.data_ptr()),
                                      C.c_void_p(stream)))
            g.status(stream)
        torch.cuda.synchronize()
        out.append((d_hits.cpu().numpy(), d_geoms.cpu().numpy(), d_cnt.cpu().numpy(), int(d_status.item())))
    return out


def _check_two_calls(g, rays, n_small, label):
    stream = torch.cuda.current_stream().cuda_stream
    d_rays = _dev(rays)
    for shadow in (False, True):
        for n in (len(rays), n_small):
            print(f"[case] {label} shadow={shadow} n={n}", flush=True)
            (h0, g0, c0, _), (h1, g1, c1, st) = _run(g, None, d_rays, n, shadow, stream)
            assert st == 0, f"{label}: status {st}"
            assert h0.tobytes() == h1.tobytes(), f"{label}: hits differ"
            assert np.array_equal(g0, g1), f"{label}: geoms differ"
            assert np.array_equal(c0, c1), f"{label}: counters differ"
    return c0.reshape(-1, 4)


def test_noalloc_equals_trace_static_scene():
    g = va.Scene(0)
    g.add_geom(va.PolyMesh(*scenes.uv_sphere()))
    g.add_geom(va.PolyMesh(*scenes.ground_quad()))
    g.pre_render()
    rays = scenes.uniform_rays(1_000_000, seed=1)
    _check_two_calls(g, rays, 123_457, "static uniform")
    # A shallow LDS stack: deeper entries spill into the scratch array, and at 32 waves per CU the
    # grid's spill columns would exceed it, so the entry point launches fewer resident waves.
    g.set_trace_config(block=64, lds_stack=8, waves_per_cu=32, sort=5)
    cnt = _check_two_calls(g, rays, 65_537, "static lds 8")
    assert cnt[:, 3].max() > 8, "no ray used the spill area"
    g.close()


def test_noalloc_equals_trace_motion_scene():
    verts, idx = scenes.moving_spheres(count=300)
    g = va.Scene(0)
    g.add_geom(va.PolyMesh(verts, idx))
    g.pre_render()
    gs = scenes.rng(3)
    n = 200_000
    rays = scenes.make_rays(gs.uniform(-25, 25, (n, 3)), scenes.uniform_sphere_dirs(gs, n), time=gs.random(n))
    _check_two_calls(g, rays, 77_777, "motion")
    g.close()


def test_noalloc_rejects_bad_calls():
    lib = _lib.load()
    g = va.Scene(0)
    g.add_geom(va.PolyMesh(*scenes.uv_sphere(1.0, 16, 8)))
    d = torch.zeros(64, dtype=torch.uint8, device="cuda")
    p = C.c_void_p(d.data_ptr())
    assert lib.vgpu_trace_noalloc(g._h, 0, p, 1, p, None, None, None, None) == _lib.VGPU_ERR_STATE
    g.pre_render()
    assert lib.vgpu_trace_noalloc(g._h, 0, p, 1 << 32, p, None, None, None, None) == _lib.VGPU_ERR_ARG
    assert b"2^32" in lib.vgpu_last_error()
    assert lib.vgpu_trace_noalloc(g._h, 0x40, p, 1, p, None, None, None, None) == _lib.VGPU_ERR_ARG
    assert lib.vgpu_trace_noalloc(g._h, 0, None, 1, p, None, None, None, None) == _lib.VGPU_ERR_ARG
    g.close()
