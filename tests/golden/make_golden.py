"""Writes tests/golden/ref_vectors.npz from the reference's own hot-path
routines (oracle/_ref, built from the unmodified Fortran sources).

Inputs are regenerated from fixed seeds by tests/golden/inputs.py, so only
outputs are stored.  Run from the repo root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.dirname(__file__)]

import inputs as I  # noqa: E402
from wxoracle import Oracle, RefLib  # noqa: E402


def main():
    out = {}
    O = Oracle()
    for variant, choice in I.FLUX_CASES:
        R = RefLib(variant)
        cfg = I.flux_cfg(variant)
        pp = cfg.params()
        m = I.flux_mesh(O, cfg)
        R.setup(pp)
        R.put(m)
        f, g = R.fluxes(1, choice)
        out[f"flux_{variant}_{choice}_f"] = f[:, 1:18, 1:18]
        out[f"flux_{variant}_{choice}_g"] = g[:, 1:18, 1:18]
    R = RefLib("c3")
    states, dts = I.atomic_inputs()
    out["atomic"] = np.stack([R.atomic(dt, s) for s, dt in zip(states, dts)])
    out["aloss"] = np.array([R.aloss(*a) for a in I.aloss_inputs()])
    R2 = RefLib("c2")
    R2.setup(I.cfg("c2").params())
    out["cooldmc"] = np.array([R2.cooldmc(T) for T in I.dmc_inputs()])
    for v in I.LIMITER_VARIANTS:
        Rl = RefLib(v)
        res = []
        for a in I.limiter_inputs():
            pl, pr = Rl.limiter(a[0], a[1], a[2], a[3])
            res.append(np.concatenate([pl, pr]))
        out[f"limiter_{v}"] = np.array(res)
    # whole steps: C1 interior after 5 steps; C2/C3 per-leaf sums after AMR steps
    for variant, nsteps in (("c1", 5), ("c2", 4), ("c3", 4)):
        R = RefLib(variant)
        cfg = I.cfg(variant)
        pp, m = I.step_mesh(O, cfg)
        R.setup(pp)
        t = 0.0
        dts = []
        for s in range(nsteps):
            O.locatebounds(pp, m)
            R.put(m)
            dt = R.L.wxr_timestep()
            t += dt
            R.L.wxr_tstep(t, dt)
            R.get(m)
            dts.append(dt)
            if cfg.nlevs > 1:
                O.admesh(pp, m)
        lv = m.leaves()
        out[f"step_{variant}_dt"] = np.array(dts)
        if variant == "c1":
            out["step_c1_u"] = m.u[lv][:, :, 2:18, 2:18]
        else:
            out[f"step_{variant}_leaves"] = np.array(lv, dtype=np.int32)
            out[f"step_{variant}_sums"] = m.u[lv][:, :, 2:18, 2:18].sum(axis=(2, 3))
    path = os.path.join(os.path.dirname(__file__), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
