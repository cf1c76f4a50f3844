"""Plain device-to-device copy of a known size, used to calibrate the
FETCH_SIZE / WRITE_SIZE hardware counters (KiB per dispatch) against bytes
actually moved, and to measure the achievable HBM bandwidth.

    python -m vermeer_amd.hbmcopy [--mib 1024] [--reps 5]
"""
import argparse
import json


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    n = a.mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    print(json.dumps({"bytes_read": n, "bytes_written": n, "ms_per_copy": ms, "GBps_read_plus_write": 2 * n / ms / 1e6}))


if __name__ == "__main__":
    main()
