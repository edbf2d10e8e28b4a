"""The device's FP64 exp / pow / log / sin (walicxe2d_amd/csrc/wx_libm.h),
compiled for the host, against the host C library on seeded random
arguments: exp, pow and sin equal to it bit for bit (the reference's rates
and losses go through exp and pow, src/coolingh.f90:41-165, the jet speed
through sin, JET/jet.f90:55), log correctly rounded (checked against
libquadmath where present); the C library's own log is not correctly
rounded on a small fraction of arguments, which the test bounds, and its sin
is not on about 1.4e-3 of them (the device's sin follows the library there
too).  Also: the committed tables equal their generator's output."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "walicxe2d_amd", "csrc")


def _cpu_has_fma_avx2():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return False
    return " fma" in flags and " avx2" in flags


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("libm") / "libm_check")
    src = os.path.join(HERE, "native", "libm_check.cpp")
    base = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", src, "-o", out]
    quad = subprocess.run(base[:-2] + ["-DWX_QUAD", "-o", out, "-lquadmath"], capture_output=True)
    if quad.returncode != 0:
        subprocess.run(base, check=True)
    return out, quad.returncode == 0


def test_tables_match_generator():
    got = subprocess.run([sys.executable, os.path.join(CSRC, "gen_libm_tables.py")], check=True,
                         capture_output=True, text=True).stdout
    assert got == open(os.path.join(CSRC, "wx_libm_tables.inc")).read()


@pytest.mark.skipif(not _cpu_has_fma_avx2(), reason="the C library takes its FMA code path only on FMA+AVX2 CPUs")
def test_libm_against_host_library(checker):
    exe, has_quad = checker
    n = 300000
    line = subprocess.run([exe, str(n)], check=True, capture_output=True, text=True).stdout.strip().splitlines()[-1]
    f = line.split()
    r = {f[k]: int(f[k + 1]) for k in range(0, len(f), 2)}
    assert r["exp"] == 0, line
    assert r["pow"] == 0, line
    assert r["sin"] == 0, line
    assert r["logpow"] == 0, line   # the shared-logarithm form of betah's log and pow
    # the C library's log differs from the correctly rounded value on ~1e-4
    # of these arguments (by one ulp)
    assert r["log"] <= 1e-3 * r["logtotal"], line
    if has_quad:
        assert r["log_cr"] == 0, line
        # the library's sin (and so the device's) is not correctly rounded
        # on a small fraction of arguments: the device does not round better
        assert 0 < r["sin_cr"] <= 5e-3 * n, line
