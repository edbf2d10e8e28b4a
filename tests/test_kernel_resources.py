"""The hot stage kernels' register budgets, as the compiler reports them
(hipcc -Rpass-analysis=kernel-resource-usage on wx_kernels.hip, gfx950):
the occupancy each was tuned for, and no scratch memory where a spill on the
hot path would reach HBM (the fill once wrote 122-138 B per zone-update
instead of 74-89 that way).  Built for the device only; no GPU needed."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def usage(tmp_path_factory):
    if not os.path.exists(HIPCC) or shutil.which("c++filt") is None:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("ru") / "k.o")
    src = os.path.join(ROOT, "walicxe2d_amd", "csrc", "wx_kernels.hip")
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "--offload-arch=gfx950",
           "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c",
           src, "-o", out, "-Rpass-analysis=kernel-resource-usage"]
    err = subprocess.run(cmd, capture_output=True, text=True, check=True).stderr
    res, name = {}, None
    for line in err.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = subprocess.run(["c++filt", m.group(1)], capture_output=True, text=True).stdout.strip()
            name = name.split("(")[0].replace("void ", "")
            res[name] = {}
            continue
        for key, pat in (("vgpr", r"VGPRs: (\d+)"), ("scratch", r"ScratchSize \[bytes/lane\]: (\d+)"),
                         ("waves", r"Occupancy \[waves/SIMD\]: (\d+)")):
            m = re.search(pat, line)
            if m and name:
                res[name][key] = int(m.group(1))
    return res


@pytest.mark.parametrize("kernel,waves,scratch", [
    ("wx::k_stageA<5, 512, false>", 8, 32),               # a few spills that stay in the caches
    ("wx::k_stageB<5, 256, false>", 4, 0),
    ("wx::k_fillI_cool<5, false, 256, false>", 5, 32),    # spills only on exp's and pow's rare paths
])
def test_stage_kernel_budgets(usage, kernel, waves, scratch):
    r = usage[kernel]
    assert r["waves"] >= waves, (kernel, r)
    assert r["scratch"] <= scratch, (kernel, r)
