"""The Fortran host path: drop-in timestep/tstep (walicxe2d_amd/fortran)
calling the library through iso_c_binding, against the CPU oracle."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from walicxe2d_amd import params as P
from wxoracle import Oracle

from harness import MASK, oracle_mesh, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/lib/llvm/bin/flang"
FLAGS = ["-cpp", "-O2", "-fdefault-real-8", "-fdefault-double-8", "-DDOUBLEP", "-DHLLC", "-DLIMITER=1",
         "-DADIABATIC", "-DOUTFXL", "-DOUTFXR", "-DOUTFYB", "-DOUTFYT"]


@pytest.mark.skipif(not os.path.exists(FC), reason="flang not installed")
def test_fortran_dropin_c1():
    nsteps = 12
    cfg = P.config_c1()
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 64)
    tmp = tempfile.mkdtemp(prefix="wxf_")
    try:
        src = os.path.join(ROOT, "walicxe2d_amd", "fortran")
        tst = os.path.join(ROOT, "tests", "fortran")
        files = [os.path.join(tst, "wxf_modules.f90"), os.path.join(src, "wx_gpu.f90"),
                 os.path.join(src, "wx_gpu_host.f90"), os.path.join(src, "tstep.f90"),
                 os.path.join(src, "timestep.f90"), os.path.join(tst, "wxf_driver.f90")]
        libdir = os.path.join(ROOT, "walicxe2d_amd")
        subprocess.run([FC, *FLAGS, *files, "-o", "wxf", f"-L{libdir}", "-lwxhip", f"-Wl,-rpath,{libdir}"],
                       cwd=tmp, check=True)
        nb = 64
        # state in the reference's layout: u(nb, ieq, i, j) with nb fastest
        uf = np.zeros((20, 20, 4, nb))            # C-order view (j, i, k, nb)
        for b in range(1, m.nbmax + 1):
            for k in range(4):
                uf[:, :, k, b - 1] = m.u[b, k]
        lpf = np.zeros((8, nb), dtype=np.int32)
        lpf[:, :m.nbmax] = m.lp[1:m.nbmax + 1].T
        leaf = np.zeros(nb, dtype=np.int32)
        leaf[:m.nbmax] = m.leaf[1:m.nbmax + 1]
        icf = np.zeros((2, nb), dtype=np.int32)
        icf[:, :m.nbmax] = m.icoord[1:m.nbmax + 1].T
        with open(os.path.join(tmp, "state_in.bin"), "wb") as fh:
            np.array([m.nbmax, m.nbleafs], dtype=np.int32).tofile(fh)
            uf.tofile(fh)
            lpf.tofile(fh)
            leaf.tofile(fh)
            icf.tofile(fh)
        subprocess.run([os.path.join(tmp, "wxf"), str(nsteps)], cwd=tmp, check=True, timeout=120)
        out = np.fromfile(os.path.join(tmp, "state_out.bin"), dtype=np.float64)
        tg, dtg = out[0], out[1]
        ug = out[2:].reshape(20, 20, 4, nb)
        t = 0.0
        for s in range(nsteps):
            dt = O.timestep(pp, m)
            t += dt
            O.tstep(pp, m, t, dt)
        assert abs(tg - t) <= 1e-12 * t and abs(dtg - dt) <= 1e-12 * dt
        lv = m.leaves()
        g = np.stack([np.stack([ug[:, :, k, b - 1] for k in range(4)]) for b in lv])
        e, where = rel_err(g, m.u[lv], MASK)
        assert e <= 1e-12, (e, where)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


AMR_BIN = os.path.join(ROOT, "oracle", "_ref", "wxf_amr_c2")


@pytest.mark.skipif(not os.path.exists(AMR_BIN), reason="oracle/_ref/wxf_amr_c2 not built (reference tree absent)")
def test_fortran_reference_host_amr_c2():
    """The reference's own host side -- its admesh (with critneighup,
    critneighdown, updatelpup, updatelpdown) and locatebounds, compiled from
    its sources into oracle/_ref/wxf_amr_c2 -- driving the GPU through the
    drop-ins tstep (host list handed over by wx_set_bounds), timestep,
    markref (device criteria), refine and coarse (device data ops), on the
    explosion with DMC cooling and 5 levels, 20 iterations of the time loop
    against the oracle's: the same tree, dt and time, and every non-corner
    cell within 1e-12."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from mesh_inputs import cfg_for
    nsteps = 20
    cfg = cfg_for("c2")        # the reference build's criteria thresholds, 4096 block IDs
    nb = 4096
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, nb - 1)
    for _ in range(3):
        assert O.admesh(pp, m) == 0
    tmp = tempfile.mkdtemp(prefix="wxf_amr_")
    try:
        neq = cfg.neq
        uf = np.zeros((20, 20, neq, nb))            # C-order view (j, i, k, nb) of u(nb, k, i, j)
        for b in range(1, m.nbmax + 1):
            for k in range(neq):
                uf[:, :, k, b - 1] = m.u[b, k]
        lpf = np.zeros((8, nb), dtype=np.int32)
        lpf[:, :m.nbmax] = m.lp[1:m.nbmax + 1].T
        leaf = np.zeros(nb, dtype=np.int32)
        leaf[:m.nbmax] = m.leaf[1:m.nbmax + 1]
        icf = np.zeros((2, nb), dtype=np.int32)
        icf[:, :m.nbmax] = m.icoord[1:m.nbmax + 1].T
        with open(os.path.join(tmp, "state_in.bin"), "wb") as fh:
            np.array([m.nbmax, m.nbleafs], dtype=np.int32).tofile(fh)
            np.array(pp.dx[:cfg.nlevs]).tofile(fh)
            np.array(pp.dy[:cfg.nlevs]).tofile(fh)
            np.array(pp.cooltab[:82]).tofile(fh)
            uf.tofile(fh)
            lpf.tofile(fh)
            leaf.tofile(fh)
            icf.tofile(fh)
        print("wxf_amr_c2:", nsteps, "steps", flush=True)
        subprocess.run([AMR_BIN, str(nsteps)], cwd=tmp, check=True, timeout=240, capture_output=True)
        with open(os.path.join(tmp, "state_out.bin"), "rb") as fh:
            tg, dtg = np.fromfile(fh, dtype=np.float64, count=2)
            ug = np.fromfile(fh, dtype=np.float64, count=20 * 20 * neq * nb).reshape(20, 20, neq, nb)
            lpg = np.fromfile(fh, dtype=np.int32, count=8 * nb).reshape(8, nb)
            leafg = np.fromfile(fh, dtype=np.int32, count=nb)
            icg = np.fromfile(fh, dtype=np.int32, count=2 * nb).reshape(2, nb)
            nbmax_g, nbl_g = np.fromfile(fh, dtype=np.int32, count=2)
        t = 0.0
        for s in range(nsteps):
            dt = O.timestep(pp, m)
            t += dt
            assert O.tstep(pp, m, t, dt) == 0
            assert O.admesh(pp, m) == 0
        assert abs(tg - t) <= 1e-12 * t and abs(dtg - dt) <= 1e-12 * dt, (tg, t, dtg, dt)
        assert (nbmax_g, nbl_g) == (m.nbmax, m.nbleafs)
        n = m.nbmax
        assert np.array_equal(lpg[:, :n].T, m.lp[1:n + 1]) and np.array_equal(icg[:, :n].T, m.icoord[1:n + 1])
        assert np.array_equal(leafg[:n], m.leaf[1:n + 1])
        lv = m.leaves()
        g = np.stack([np.stack([ug[:, :, k, b - 1] for k in range(neq)]) for b in lv])
        e, where = rel_err(g, m.u[lv], MASK)
        assert e <= 1e-12, (e, where)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
