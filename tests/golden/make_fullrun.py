"""Writes tests/golden/fullrun_<workload>.npz: the oracle's whole runs of the
explosion with AMR and DMC cooling (8x8 roots, 5 levels, to 20 kyr) and of
the cooling jet (64x16 roots, 4 levels, to 300 yr), from the initial state
through the 4 initial admesh passes (src/main.f90:60-63) and the time loop
(src/main.f90:87-98: timestep, time += dt, tstep, admesh) until the time
passes the end time.  Stored: the number of steps and the dt of every step,
a SHA-256 digest of the tree and the leaves' conserved variables on the
non-corner cells (fullrun_common.digest_u) every 100 steps and at the end,
the digest of tests/golden/mesh_inputs.py (primitives too) at the end, the
totals of every
conserved variable over the leaves (cell area weighted) at the start and at
the end, and the final density and pressure projected to the level-1 grid
(mean over each root block's finest cells).

    python tests/golden/make_fullrun.py c2|c3     (from the repo root; the
    jet takes about two hours of one core)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.dirname(__file__)]

import fullrun_common as FC  # noqa: E402
import mesh_inputs as MI  # noqa: E402
from harness import oracle_mesh  # noqa: E402
from wxoracle import Oracle  # noqa: E402


def main(name):
    cfg, cap, tend = FC.workload(name)
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, cap)
    for _ in range(FC.INITIAL_ADMESH):
        assert O.admesh(pp, m) == 0
    tot0 = FC.totals(cfg, m.u, m.lp, m.leaf)
    t = 0.0
    dts, digs, at = [], [], []
    t0 = time.time()
    s = 0
    while t <= tend:
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        assert O.admesh(pp, m) == 0
        dts.append(dt)
        s += 1
        if s % FC.DIGEST_EVERY == 0:
            digs.append(FC.digest_u(m.nbmax, m.nbleafs, m.lp, m.leaf, m.icoord, m.u))
            at.append(s)
            print(name, s, t, m.nbleafs, "%.0fs" % (time.time() - t0), flush=True)
    digs.append(FC.digest_u(m.nbmax, m.nbleafs, m.lp, m.leaf, m.icoord, m.u))
    at.append(s)
    rho1, p1 = FC.level1_fields(O, pp, cfg, m)
    out = dict(nsteps=np.int64(s), tend=np.float64(tend), time=np.float64(t), dt=np.array(dts),
               digest_u=np.stack(digs), digest_step=np.array(at, dtype=np.int64),
               digest_final=MI.state_digest(m),
               totals0=tot0, totals=FC.totals(cfg, m.u, m.lp, m.leaf), rho1=rho1, p1=p1,
               nbleafs=np.int64(m.nbleafs), nbmax=np.int64(m.nbmax))
    path = os.path.join(os.path.dirname(__file__), f"fullrun_{name}.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes", s, "steps", flush=True)
    # the final state itself, for diagnostics (not a fixture: too large)
    keep = os.environ.get("WX_FULLRUN_STATE")
    if keep:
        lv = m.leaves()
        np.savez_compressed(keep, u=m.u[:m.nbmax + 1], lp=m.lp[:m.nbmax + 1], leaf=m.leaf[:m.nbmax + 1],
                            icoord=m.icoord[:m.nbmax + 1], nbmax=m.nbmax, nbleafs=m.nbleafs, leaves=np.array(lv))


if __name__ == "__main__":
    main(sys.argv[1])
