"""GPU path (libwxhip.so through the C-ABI) against the CPU oracle."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Simulation
from wxoracle import Oracle

from harness import MASK, MASK_INTERIOR, oracle_mesh, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12   # relative, per cell per step


def _pair_run(cfg, nsteps, cap, admesh, tol=TOL, label=""):
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, cap)
    sim = Simulation(cfg, nblocks=cap)
    sim.basegrid()
    if admesh and cfg.nlevs > 1:
        for _ in range(3):
            O.admesh(pp, m)
            sim.admesh()
    t = 0.0
    worst = 0.0
    for s in range(nsteps):
        lp, leaf, ic, nbmax, nbl = sim.tree()
        assert nbmax == m.nbmax and nbl == m.nbleafs, (s, nbmax, m.nbmax)
        assert (lp[1:] == m.lp[1:nbmax + 1]).all()
        assert (leaf[1:] == m.leaf[1:nbmax + 1]).all()
        assert (ic[1:] == m.icoord[1:nbmax + 1]).all()
        dt_o = O.timestep(pp, m)
        t += dt_o
        assert O.tstep(pp, m, t, dt_o) == 0
        tg, dtg, _ = sim.step(1, admesh=False)
        assert abs(dtg - dt_o) <= TOL * dt_o, (s, dtg, dt_o)
        assert abs(tg - t) <= TOL * t
        ug = sim.download()
        lv = m.leaves()
        e, where = rel_err(ug[lv], m.u[lv], MASK)
        worst = max(worst, e)
        assert e <= tol, f"{label} step {s}: rel err {e:.3e} at var/leaf/cell {where}"
        if admesh and cfg.nlevs > 1:
            assert O.admesh(pp, m) == 0
            sim.admesh()
    sim.close()
    return worst


def test_c1_explosion_steps():
    """C1 (adiabatic, single level, outflow, OUTFYB compat) for 40 steps."""
    cfg = P.config_c1()
    w = _pair_run(cfg, 40, 64, admesh=False, label="c1")
    print("c1 worst rel err", w)
