"""Seeded inputs shared by tests/golden/make_golden.py (which stores the
reference's outputs) and tests/test_oracle_golden.py (which checks the
oracle against them).  numpy's PCG64 streams are platform independent."""
import numpy as np

from walicxe2d_amd import params as P
from wxoracle import MeshState

FLUX_CASES = [("c1", 1), ("c1", 2), ("hll", 2), ("hybrid", 2), ("lim2", 2), ("cyl", 2)]
LIMITER_VARIANTS = ["lim0", "c1", "lim2", "lim3", "lim4", "lim5", "lim6"]


def cfg(variant):
    if variant == "c1":
        return P.config_c1()
    if variant == "c2":
        return P.config_c2(nblocks=4096)
    if variant == "c3":
        return P.config_c3(nblocks=4096, nbrootx=16, nbrooty=4)
    raise KeyError(variant)


def flux_cfg(variant):
    if variant in ("c1",):
        return cfg("c1")
    base = dict(name=variant, neq=5, nlevs=3, nbrootx=4, nbrooty=4, xmax=1.0, ymax=1.0, xphys=P.PC,
                amol=0.6, cooling="ADIABATIC", bc=("REFLECTIVE", "OUTFLOW", "OUTFLOW", "REFLECTIVE"),
                ic="explosion_code", ic_args=dict(rho=1.0, p=1e-2, E=1.0, r=0.1, xc=0.4, yc=0.45))
    if variant == "hll":
        base["solver"] = "HLL"
    elif variant == "hybrid":
        base["solver"] = "HLL_HLLC"
    elif variant.startswith("lim"):
        base["limiter"] = int(variant[3:])
    elif variant == "cyl":
        base["limiter"] = 2
        base["cylindrical"] = True
    return P.Config(**base)


def flux_mesh(O, c):
    """Block 1 holds seeded random primitives (both Riemann branches occur)."""
    m = MeshState(32, c.neq)
    O.basegrid(c.params(), m)
    rng = np.random.default_rng(11)
    q = np.empty((c.neq, 20, 20))
    q[0] = 10 ** rng.uniform(-1, 1, (20, 20))
    q[1] = rng.uniform(-2, 1, (20, 20))
    q[2] = rng.uniform(-1, 1, (20, 20))
    q[3] = 10 ** rng.uniform(-2, 0, (20, 20))
    for k in range(4, c.neq):
        q[k] = q[0] * rng.uniform(0, 1, (20, 20))
    m.primit[1] = q
    return m


def atomic_inputs():
    rng = np.random.default_rng(12)
    pp = cfg("c3").params()
    states, dts = [], []
    for _ in range(300):
        n = 10 ** rng.uniform(-1, 4)
        T = 10 ** rng.uniform(1, 8)
        y = rng.uniform(0, 1)
        vx, vy = rng.normal(0, 1, 2)
        Pth = n * (2 - y) * T / pp.Tempsc
        states.append(np.array([n, n * vx, n * vy, 1.5 * Pth + 0.5 * n * (vx * vx + vy * vy), n * y]))
        dts.append(10 ** rng.uniform(2, 11))
    return states, dts


def aloss_inputs():
    rng = np.random.default_rng(13)
    return [(rng.uniform(0, 1), rng.uniform(0, 1), 10 ** rng.uniform(2, 11), 10 ** rng.uniform(-1, 4),
             10 ** rng.uniform(-2, 3), 10 ** rng.uniform(2, 8)) for _ in range(300)]


def dmc_inputs():
    return np.concatenate([10 ** np.linspace(4.0003, 7.9997, 300), [1.5e8, 3e9]])


def limiter_inputs():
    rng = np.random.default_rng(14)
    a = rng.normal(0, 1, (200, 4, 5))
    a[::7, :, 2] = 0.0
    return a


def step_mesh(O, c):
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    from harness import oracle_mesh
    pp, m = oracle_mesh(O, c, 3000)
    if c.nlevs > 1:
        for _ in range(3):
            O.admesh(pp, m)
    return pp, m
