"""Writes tests/golden/ref_mesh_vectors.npz from the reference's own mesh
layer and time loop (oracle/_ref, built from the unmodified Fortran sources):
the locatebounds list (in order) of each seeded random mesh, and for each
whole-loop case the dt of every step and a SHA-256 digest of the tree and
the leaves' state after every step (tests/golden/mesh_inputs.py).

Run from the repo root: python tests/golden/make_golden_mesh.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.dirname(__file__)]

import mesh_inputs as MI  # noqa: E402
from harness import oracle_mesh  # noqa: E402
from wxoracle import Oracle, RefLib  # noqa: E402


def main():
    out = {}
    O = Oracle()
    libs = {}
    for variant, seed in MI.LB_CASES:
        R = libs.setdefault(variant, RefLib(variant))
        cfg, pp, m = MI.random_mesh(O, variant, seed)
        R.setup(pp)
        R.put(m)
        out[f"lb_{variant}_{seed}"] = R.locatebounds(0.0)
    for variant, steps in MI.LOOP_CASES:
        R = libs.setdefault(variant, RefLib(variant))
        cfg = MI.cfg_for(variant)
        pp, m = oracle_mesh(O, cfg, 3000)
        R.setup(pp)
        R.put(m)
        if cfg.nlevs > 1:
            for _ in range(3):
                R.admesh()
        t = 0.0
        dts, dig = [], []
        for s in range(steps):
            dt = R.L.wxr_timestep()
            t += dt
            R.L.wxr_tstep(t, dt)
            if cfg.nlevs > 1:
                R.admesh()
            R.get(m)
            dts.append(dt)
            dig.append(MI.state_digest(m))
        out[f"loop_{variant}_dt"] = np.array(dts)
        out[f"loop_{variant}_digest"] = np.stack(dig)
    path = os.path.join(os.path.dirname(__file__), "ref_mesh_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
