"""Achievable HBM bandwidth and counter calibration with the library's streaming kernels
(vgpu_diag_stream: 16-byte nontemporal loads and stores, four in flight per thread).

A buffer far larger than the L2 and the MALL is copied (read + write) and read; the best rate over a
few grid sizes is reported.  Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE the same kernels move a known
number of bytes, which gives the scale factor from those counters to bytes.

    python -m vermeer_amd.hbmcopy [--mib 4096] [--reps 5]
"""
import argparse
import ctypes as C
import json

from . import _lib


def measure(torch, mib=4096, reps=5, grids=(4, 8, 16, 32)):
    """Returns {"copy_GBps": read+write GB/s, "read_GBps": GB/s, per-grid rates, sizes}."""
    lib = _lib.load()
    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    sums = torch.empty(cus * max(grids) * 256, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {"bytes": n, "copy": {}, "read": {}}
    for mode, name, d in ((_lib.VGPU_DIAG_COPY, "copy", dst), (_lib.VGPU_DIAG_READ, "read", sums)):
        for bpc in grids:
            def run():
                _lib.check(lib.vgpu_diag_stream(mode, C.c_void_p(src.data_ptr()), C.c_void_p(d.data_ptr()), n, bpc,
                                                C.c_void_p(st)))
            run()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / reps
            moved = 2 * n if mode == _lib.VGPU_DIAG_COPY else n
            out[name][bpc] = round(moved / ms / 1e6, 1)
    out["copy_GBps"] = max(out["copy"].values())
    out["read_GBps"] = max(out["read"].values())
    del src, dst, sums
    return out


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    print(json.dumps(measure(torch, a.mib, a.reps)))


if __name__ == "__main__":
    main()
