"""Whole runs on the GPU against the oracle: the explosion (C1) for its 200
steps stepped beside the oracle, and the AMR explosion with DMC cooling (to
20 kyr) and the cooling jet (to 300 yr) against fixtures of the oracle's
whole runs (tests/golden/make_fullrun.py): dt of every step, a digest of the
tree and the leaves' state every 100 steps (printed every 1000) and at the
end, and at the end
the relative L1 difference of density and pressure on the level-1 grid
(held to L1_TOL) and the drift of the conserved totals."""
import os
import sys

import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Simulation
from wxoracle import Oracle

from harness import MASK, oracle_mesh, rel_err

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import fullrun_common as FC  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# relative L1 of density and pressure on the level-1 grid after a whole run,
# the bound for a run that is not bit-identical: the device's exp, pow and
# sin are the host library's and its log the correctly rounded one (one ulp
# from the library's on ~2e-4 of arguments, in the recombination loss only);
# the jet's shear layers amplify any difference over the run
L1_TOL = 1e-6


def test_c1_full_run_bitexact():
    """C1 (adiabatic explosion, OUTFYB compat) for its 200 steps: the device
    runs its own time loop beside the oracle's, never re-synchronised; dt
    equal every step and every non-corner cell bit-identical at the end."""
    cfg = P.config_c1()
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 64)
    sim = Simulation(cfg, nblocks=64)
    sim.basegrid()
    t = 0.0
    for s in range(200):
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        tg, dtg, _ = sim.step(1, admesh=False)
        assert dtg == dt and tg == t, (s, dtg, dt)
    lv = m.leaves()
    ug = sim.download()
    sim.close()
    assert np.array_equal(ug[lv][:, :, MASK], m.u[lv][:, :, MASK])


def _full_run(name):
    path = os.path.join(GOLD, f"fullrun_{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"{path} not written (tests/golden/make_fullrun.py {name})")
    fx = np.load(path)
    cfg, cap, tend = FC.workload(name)
    sim = Simulation(cfg, nblocks=cap)
    sim.basegrid()
    for _ in range(FC.INITIAL_ADMESH):
        sim.admesh()
    lp, leaf, ic, nbmax, nbl = sim.tree()
    tot0 = FC.totals(cfg, sim.download(), lp, leaf)
    assert np.array_equal(tot0, fx["totals0"])
    dts = fx["dt"]
    at = list(fx["digest_step"])
    first_diff = None
    t = 0.0
    for s in range(int(fx["nsteps"])):
        tg, dtg, _ = sim.step(1, admesh=True)
        t += dts[s]
        if first_diff is None and dtg != dts[s]:
            first_diff = ("dt", s + 1)
        k = at.index(s + 1) if (s + 1) in at else -1
        if k >= 0 and (first_diff is None or k % 10 == 9 or k == len(at) - 1):
            lp, leaf, ic, nbmax, nbl = sim.tree()
            d = FC.digest_u(nbmax, nbl, lp, leaf, ic, sim.download())
            same = np.array_equal(d, fx["digest_u"][k])
            if k % 10 == 9 or k == len(at) - 1 or (first_diff is None and not same):
                print(f"{name} step {s + 1}: digest {'equal' if same else 'differs'}", flush=True)
            if first_diff is None and not same:
                first_diff = ("state", s + 1)
    assert abs(tg - float(fx["time"])) <= 1e-12 * float(fx["time"])
    u, prim = sim.download(prim=True)
    lp, leaf, ic, nbmax, nbl = sim.tree()
    sim.close()
    rho1, p1 = FC.level1(cfg, u, prim, lp, leaf, ic)
    l1r = float(np.abs(rho1 - fx["rho1"]).sum() / np.abs(fx["rho1"]).sum())
    l1p = float(np.abs(p1 - fx["p1"]).sum() / np.abs(fx["p1"]).sum())
    tot = FC.totals(cfg, u, lp, leaf)
    drift = np.abs(tot - fx["totals"]) / np.maximum(np.abs(fx["totals"]), 1e-300)
    print(f"{name}: {int(fx['nsteps'])} steps, first difference {first_diff}, L1 rho {l1r:.3e} P {l1p:.3e}, "
          f"totals rel. diff {drift.max():.3e}", flush=True)
    assert l1r <= L1_TOL and l1p <= L1_TOL, (l1r, l1p, first_diff)
    return first_diff, l1r, l1p


def test_c2_full_run():
    """The AMR explosion with DMC cooling (8x8 roots, 5 levels) to 20 kyr,
    the cooled pressure carried from each step into the next as the
    reference carries primit: the digests equal the oracle's at every check
    and L1 is zero."""
    first_diff, l1r, l1p = _full_run("c2")
    assert first_diff is None and l1r == 0.0 and l1p == 0.0, (first_diff, l1r, l1p)


def test_c3_full_run():
    """The cooling jet (64x16 roots, 4 levels, coolingh, jet source) to 300
    yr; the digests equal the oracle's at every check and L1 is zero."""
    first_diff, l1r, l1p = _full_run("c3")
    assert first_diff is None and l1r == 0.0 and l1p == 0.0, (first_diff, l1r, l1p)
