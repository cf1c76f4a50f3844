"""Runs one workload a few times on device-resident inputs, for rocprofv3 (kernel trace, counters):

    python -m vermeer_amd.profile_workloads sphere_ground   # 65,026 triangles: 1024^2 pinhole and 1M uniform
                                                            # rays, closest and any hit
    python -m vermeer_amd.profile_workloads moving_spheres  # 500,000 moving triangles: 1024^2 x 4 camera rays
                                                            # and their point-light shadow rays
    python -m vermeer_amd.profile_workloads room            # 1,004,048-triangle room, plain path mode, 1024^2
    python -m vermeer_amd.profile_workloads stream          # 1 GiB streaming copy and read (counter calibration)

Each launch set runs once untimed, then --reps times; progress goes to stderr.
"""
import argparse
import ctypes as C
import sys

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _launches(torch, sc, d_rays, n, shadow, reps, label):
    st = torch.cuda.current_stream().cuda_stream
    h = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
    for _ in range(reps + 1):
        sc.trace_device(d_rays.data_ptr(), n, h.data_ptr(), shadow=shadow, stream=st)
    sc.status(st)
    log(f"[profile] {label}: {reps + 1} launches of {n} rays")
    return h


def sphere_ground(torch, reps):
    import vermeer_amd as va
    from vermeer_amd import scenes
    sc = va.Scene(0)
    sc.add_geom(va.PolyMesh(*scenes.uv_sphere()))
    sc.add_geom(va.PolyMesh(*scenes.ground_quad()))
    sc.pre_render()
    for name, rays in (("pinhole 1024^2", scenes.pinhole_rays()), ("uniform 1M", scenes.uniform_rays())):
        d = _dev(torch, rays)
        for shadow in (False, True):
            _launches(torch, sc, d, len(rays), shadow, reps, f"{name} {'any hit' if shadow else 'closest'}")
    sc.close()


def moving_spheres(torch, reps):
    import vermeer_amd as va
    from vermeer_amd import scenes
    verts, idx = scenes.moving_spheres()
    sc = va.Scene(0)
    sc.add_geom(va.PolyMesh(verts, idx))
    sc.pre_render()
    rays = scenes.moving_spheres_camera_rays()
    n = len(rays)
    h = _launches(torch, sc, _dev(torch, rays), n, False, reps, "camera rays closest")
    hits = h.cpu().numpy().view(va.HIT_DTYPE)
    srays = scenes.point_light_shadow_rays(rays, hits["prim"], hits["t"])
    _launches(torch, sc, _dev(torch, srays), n, True, reps, "shadow rays any hit")
    sc.close()


def room(torch, reps):
    from vermeer_amd import pathtrace, scenes
    desc = scenes.guided_room_scene()
    W = 1024
    pt = pathtrace.PathTracer(desc, W, W, max_depth=8, seed=1)
    acc = torch.zeros(W * W * 3, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for p in range(reps + 1):
        pt.render_pass_device(acc.data_ptr(), p, stream=st)
    torch.cuda.synchronize()
    log(f"[profile] room: {reps + 1} plain passes at {W}^2, depth 8")
    pt.close()


def stream(torch, reps):
    from vermeer_amd import _lib
    lib = _lib.load()
    n = 1 << 30
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    sums = torch.empty(cus * 8 * 256, dtype=torch.float32, device="cuda")
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for mode, d in ((_lib.VGPU_DIAG_COPY, dst), (_lib.VGPU_DIAG_READ, sums)):
        for _ in range(reps + 1):
            _lib.check(lib.vgpu_diag_stream(mode, C.c_void_p(src.data_ptr()), C.c_void_p(d.data_ptr()), n, 8, st))
    torch.cuda.synchronize()
    log(f"[profile] stream: {reps + 1} copies and reads of 1 GiB")


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["sphere_ground", "moving_spheres", "room", "stream"])
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    {"sphere_ground": sphere_ground, "moving_spheres": moving_spheres, "room": room, "stream": stream}[a.workload](
        torch, a.reps)


if __name__ == "__main__":
    main()
