"""Builds the gfx950 shared library walicxe2d_amd/libwxhip.so in-tree.

    python -m walicxe2d_amd.build        (or __graft_entry__.build())

hipcc cross-compiles for gfx950 without a GPU present.  Floating-point
contraction is off so the FP64 arithmetic keeps the reference's operation
order; packed FP32 instructions are disabled for every HIP translation unit.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwxhip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function",
]
DEVICE = [
    "--offload-arch=gfx950", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# device translation units
KERNEL_SOURCES = ["wx_kernels.hip", "wx_surrogate.hip", "wx_noalloc.hip", "wx_lbuild.hip"]
# host-side translation units
HOST_SOURCES = ["wx_mesh.cpp", "wx_runtime.cpp", "wx_host.cpp", "wx_ensemble.cpp", "wx_surrogate.cpp", "wx_bin.cpp",
                "wx_loadbalance.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for name in KERNEL_SOURCES:
        jobs.append(([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, name)], os.path.join(objdir, name + ".o")))
    for name in HOST_SOURCES:
        jobs.append(([HIPCC, *COMMON, *DEVICE, "-x", "hip", "-c", os.path.join(CSRC, name)],
                     os.path.join(objdir, name + ".o")))
    # one compiler process per translation unit, a few at a time
    workers = max(1, min(len(jobs), 8, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd + ["-o", obj]) for cmd, obj in jobs]:
            f.result()
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
