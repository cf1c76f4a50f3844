#!/usr/bin/env python
"""Incoherent closest-hit traversal throughput on one MI355X.

Workload (the largest traversal configuration that fits one GPU): a field of
4096 icospheres (subdiv 3, 1280 triangles each, 5,242,880 triangles,
radii log-uniform in [0.01,0.5], centres uniform in [0,100]^3) and
16,777,216 cosine-weighted secondary rays leaving random surface points, each
tagged with a path depth drawn from U{1..8}.  Rays are compacted by depth
(a stable partition) and one traversal launch per bounce processes that
bounce's rays; a step is the 8 launches over all 16,777,216 rays.

Inputs are resident in HBM before the timed region.  The timed region is K
steps bracketed by torch.cuda.synchronize(); HIP events around every launch
give the traversal kernel's average duration, from which the roofline's
achieved bandwidth is computed with the algorithmic byte model
    bytes/ray = 32 (ray) + 16 (hit) + 128 * node visits + 48 * triangle tests
using per-ray visit counts taken from one counting pass of the same kernel
(they equal the oracle's counts, which the GPU tests check).

The CPU baseline is the oracle (a C++ restatement of the reference's Go
traversal, OpenMP over the host cores) on a bounded sample of the same rays.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s closest-hit (incoherent); guided vs plain time-to-equal-RMSE on 1 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--spheres", type=int, default=4096)
    ap.add_argument("--trace-config", default="", help="block,lds_stack,waves_per_cu (default: autotune)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work for the baseline sample")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph replay measurement")
    return ap.parse_args()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU traversal shards rays across GPUs with a replicated BVH and is not part of this build",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import vermeer_amd as va
    from vermeer_amd import scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")

    t0 = time.time()
    verts, idx = scenes.icosphere_field(count=args.spheres)[:2]
    d_rays_unsorted, d_depth = scenes.bounce_rays_device(verts, idx, args.rays, seed=17, device=dev)
    # compaction by bounce: stable partition by path depth, one contiguous segment per bounce
    depth_sorted, order = torch.sort(d_depth, stable=True)
    d_rays = d_rays_unsorted.view(-1, 32)[order].reshape(-1).contiguous()
    del d_rays_unsorted, order
    seg_counts = torch.bincount(depth_sorted.long(), minlength=9)[1:9].cpu().numpy()
    seg_start = np.concatenate([[0], np.cumsum(seg_counts)[:-1]])
    torch.cuda.synchronize()
    log(f"[bench] generated {len(idx)} triangles, {args.rays} rays in {time.time() - t0:.1f}s; per bounce {seg_counts.tolist()}")

    t0 = time.time()
    sc = va.Scene(0)
    sc.add_geom(va.PolyMesh(verts, idx))
    sc.pre_render()
    build_s = time.time() - t0
    info = sc.info()
    log(f"[bench] host BVH build + upload {build_s:.2f}s, {int(info['geom_nodes'][0])} nodes")

    n = args.rays
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    d_cnt = torch.empty(n * 4, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    st = stream.cuda_stream
    rp = d_rays.data_ptr()
    hp = d_hits.data_ptr()

    if args.trace_config:
        b, l, w = (int(x) for x in args.trace_config.split(","))
        sc.set_trace_config(b, l, w)
        tuned_ms = None
    else:
        cfg, tuned_ms = sc.autotune(rp, n, hp, stream=st)
        log(f"[bench] autotuned trace config {cfg} ({tuned_ms:.2f} ms per full pass)")
    cfg = sc.trace_config()

    # counting pass (untimed): per-ray node visits and triangle tests for the byte model
    sc.trace_device(rp, n, hp, d_counters=d_cnt.data_ptr(), stream=st)
    sc.status(st)
    cnt = d_cnt.view(n, 4).to(torch.float64)
    mean_nodes = float(cnt[:, 0].mean())
    mean_tris = float(cnt[:, 2].mean())
    max_depth = int(d_cnt.view(n, 4)[:, 3].max())
    bytes_per_ray = 32.0 + 16.0 + 128.0 * mean_nodes + 48.0 * mean_tris
    del cnt
    log(f"[bench] mean node visits {mean_nodes:.2f}, triangle tests {mean_tris:.2f}, max stack {max_depth}, "
        f"{bytes_per_ray:.0f} algorithmic B/ray")

    segs = [(int(s), int(c)) for s, c in zip(seg_start, seg_counts) if c > 0]

    def step(evs=None):
        for j, (s, c) in enumerate(segs):
            if evs is not None:
                evs[j][0].record(stream)
            sc.trace_device(rp + s * 32, c, hp + s * 16, stream=st)
            if evs is not None:
                evs[j][1].record(stream)

    for _ in range(args.warmup):
        step()
    sc.status(st)

    K = args.steps
    evs = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)] for _ in segs]
           for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sc.status(st)
    kern_ms = [sum(e[0].elapsed_time(e[1]) for e in evs[k]) for k in range(K)]
    launch_ms = [e[0].elapsed_time(e[1]) for k in range(K) for e in evs[k]]
    ms_per_step = wall * 1000.0 / K
    kernel_ms_per_step = float(np.mean(kern_ms))
    mrays = n * K / wall / 1e6
    achieved = bytes_per_ray * n / (kernel_ms_per_step * 1e-3) / 1e9
    log(f"[bench] {ms_per_step:.3f} ms/step wall, {kernel_ms_per_step:.3f} ms kernel, {mrays:.1f} Mrays/s, "
        f"{achieved:.0f} GB/s algorithmic")

    graph = None
    if not args.no_graph:
        # the bounce loop captured once and replayed: the 8 launches as one hipGraph
        g = torch.cuda.CUDAGraph()
        s2 = torch.cuda.Stream()
        s2.wait_stream(stream)
        with torch.cuda.stream(s2):
            st2 = s2.cuda_stream
            for s, c in segs:   # warm on the capture stream (spill area already allocated)
                sc.trace_device(rp + s * 32, c, hp + s * 16, stream=st2)
            s2.synchronize()
            with torch.cuda.graph(g, stream=s2):
                for s, c in segs:
                    sc.trace_device(rp + s * 32, c, hp + s * 16, stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        ref = d_hits.clone()
        for _ in range(2):
            g.replay()
        torch.cuda.synchronize()
        assert torch.equal(ref, d_hits), "graph replay differs from eager launches"
        t0 = time.perf_counter()
        for _ in range(K):
            g.replay()
        torch.cuda.synchronize()
        gwall = time.perf_counter() - t0
        graph = {"ms_per_step": gwall * 1000.0 / K, "mrays_per_s": n * K / gwall / 1e6}
        log(f"[bench] hipGraph replay {graph['ms_per_step']:.3f} ms/step, {graph['mrays_per_s']:.1f} Mrays/s")

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = oracle.OracleScene()
        o.add_mesh(verts, idx)
        o.commit()
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = min(threads, 16)
        probe = 20000
        t0 = time.perf_counter()
        o.trace(d_rays[: probe * 32].cpu().numpy().view(va.RAY_DTYPE), nthreads=threads)
        dt = time.perf_counter() - t0
        m = int(min(n, max(probe, probe * args.cpu_seconds / max(dt, 1e-6))))
        pick = torch.linspace(0, n - 1, m, device=dev).long()
        sample = d_rays.view(-1, 32)[pick].cpu().numpy().view(va.RAY_DTYPE).reshape(-1)
        t0 = time.perf_counter()
        o.trace(sample, nthreads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": round(m / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{m} of the {n} bounce rays (evenly strided across all bounces), closest hit, "
                         f"C++ restatement of the Go traversal, OpenMP {threads} threads, {dt:.1f}s"}
        log(f"[bench] CPU baseline {cpu['value']} Mrays/s on {threads} threads")

    # HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs (calibrated against a
    # plain copy), scaled to this run's mean rays per launch; null when no measurement matches this kernel
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_trace_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        if tj.get("workload") == "icosphere_field_bounce_rays" and tj["kernel"].startswith(
                f"k_trace<{cfg['block']},{cfg['lds_stack']},"):
            traffic = round(tj["hbm_bytes_per_ray"] * n / len(segs))

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "icosphere_field_bounce_rays",
            "triangles": int(len(idx)),
            "rays_per_step": int(n),
            "launches_per_step": len(segs),
            "rays_per_bounce": [int(c) for c in seg_counts],
            "trace_config": cfg,
            "bvh_nodes": int(info["geom_nodes"][0]),
            "mean_node_visits": round(mean_nodes, 3),
            "mean_triangle_tests": round(mean_tris, 3),
            "max_stack_depth": max_depth,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_ray": round(bytes_per_ray, 1),
            "kernel_ms_per_step": round(kernel_ms_per_step, 4),
            "kernel_avg_launch_ms": round(float(np.mean(launch_ms)), 4),
        },
        "cpu_baseline": cpu,
        "hipgraph": graph,
        "host_build_s": round(build_s, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
