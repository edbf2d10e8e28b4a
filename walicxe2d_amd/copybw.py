"""Streaming-copy bandwidth of the GPU (the roofline's measured HBM ceiling)
and the calibration kernel for rocprofv3 FETCH_SIZE / WRITE_SIZE counters:
one device-to-device copy of a known number of bytes per dispatch.

    python -m walicxe2d_amd.copybw [--mib 1024] [--reps 20]
"""
from __future__ import annotations

import argparse
import json
import time


def main():
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    n = a.mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nbytes = 2 * n * 8
    print(json.dumps({"copy_bytes_per_dispatch_read": n * 8, "copy_bytes_per_dispatch_write": n * 8,
                      "GBps_read_plus_write": nbytes * a.reps / el / 1e9, "reps": a.reps}))


if __name__ == "__main__":
    main()
