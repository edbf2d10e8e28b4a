"""States of the BIN-writer comparison with the reference's own output
(src/output.f90:39-67), regenerated from fixed seeds with the oracle and
shared by tests/test_binio.py and tests/golden/make_golden_bin.py; the
reference's files are kept as SHA-256 digests (ref_bin_digests.json)."""
import hashlib
import os

import numpy as np

import mesh_inputs as MI
from harness import oracle_mesh

# (variant of oracle/_ref, seed of the random mesh or None for the initial
# roots, oracle steps taken before writing)
BIN_CASES = [("c1", None, 5), ("hll", 3, 2), ("c3", 1, 2)]
REF_NBLOCKS = 4096          # block capacity the reference build is compiled with
FLANG_TRUE = 1              # a .true. logical as flang stores it


def state(O, variant, seed, steps):
    if seed is None:
        cfg = MI.cfg_for(variant)
        pp, m = oracle_mesh(O, cfg, 3000)
    else:
        cfg, pp, m = MI.random_mesh(O, variant, seed)
    t = 0.0
    for _ in range(steps):
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
    return cfg, pp, m


def write_state_file(path, m, cfg, ltrue=FLANG_TRUE, nblocks=REF_NBLOCKS):
    """The state file of tests/hostsim's writebin mode."""
    n = m.nbmax
    with open(path, "wb") as f:
        np.array([cfg.neq, nblocks, cfg.nlevs, n, m.nbleafs, ltrue], dtype=np.int32).tofile(f)
        np.ascontiguousarray(m.lp[1:n + 1], dtype=np.int32).tofile(f)
        np.ascontiguousarray(m.leaf[1:n + 1], dtype=np.int32).tofile(f)
        np.ascontiguousarray(m.icoord[1:n + 1], dtype=np.int32).tofile(f)
        np.ascontiguousarray(m.u[:n + 1], dtype=np.float64).tofile(f)


def reference_files(R, pp, m, outdir):
    """The reference's output(1) of state m (its OUTBIN pair, rank 0) under
    outdir/BIN/."""
    os.makedirs(os.path.join(outdir, "BIN"), exist_ok=True)
    R.setup(pp)
    R.put(m)
    cwd = os.getcwd()
    os.chdir(outdir + os.sep)
    try:
        R.L.wxr_output(1)
    finally:
        os.chdir(cwd)
    return (os.path.join(outdir, "BIN", "pointers000.001.bin"), os.path.join(outdir, "BIN", "blocks000.001.bin"))


def digest(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()
