"""Shared by tests/golden/make_fullrun.py and the full-run GPU test: the two
full workloads, their end times, and the diagnostics kept of a run."""
import hashlib

import numpy as np

from walicxe2d_amd import params as P

INITIAL_ADMESH = 4      # src/main.f90:60-63
DIGEST_EVERY = 100


def workload(name):
    """(config, block capacity, end time in code units)."""
    if name == "c2":
        cfg = P.config_c2(nblocks=100000)
        tend_s = 20.e3 * P.YR
    elif name == "c3":
        cfg = P.config_c3(nblocks=12000)
        tend_s = 300. * P.YR
    else:
        raise ValueError(name)
    return cfg, cfg.nblocks, tend_s / cfg.params().tsc


def cell_areas(cfg, lp, leaf):
    """Cell area (code units) of every block ID's level, 0 for non-leaves."""
    pp = cfg.params()
    a = np.zeros(len(leaf))
    for nb in np.nonzero(leaf)[0]:
        lev = int(lp[nb, 0])
        a[nb] = pp.dx[lev - 1] * pp.dy[lev - 1]
    return a


def totals(cfg, u, lp, leaf):
    """Sum over the leaves' interior cells of each conserved variable times
    the cell area (Cartesian; mass, momenta, energy, scalars)."""
    a = cell_areas(cfg, lp, leaf)
    s = u[:, :, 2:18, 2:18].sum(axis=(2, 3))          # [block][var]
    return (s * a[:, None]).sum(axis=0)


def level1(cfg, u, prim, lp, leaf, icoord):
    """Density and pressure on the level-1 grid: each level-1 cell is the
    mean over the finest cells it covers of the leaves' values (the L1
    projection of walicxe2d_amd/l1.py, then averaged 2^(nlevs-1) x
    2^(nlevs-1)).  u, prim: tiles by block ID."""
    from walicxe2d_amd import l1
    f = 2 ** (cfg.nlevs - 1)
    out = []
    for arr, var in ((u, 0), (prim, 3)):
        g = l1.project(cfg, arr, lp, leaf, icoord, var=var)
        ny, nx = g.shape
        out.append(g.reshape(ny // f, f, nx // f, f).mean(axis=(1, 3)))
    return out[0], out[1]


def level1_fields(O, pp, cfg, m):
    """level1() of an oracle state (primitives recomputed by calcprim)."""
    prim = np.zeros_like(m.u)
    for nb in m.leaves():
        prim[nb] = O.calcprim(pp, m.u[nb])
    return level1(cfg, m.u, prim, m.lp, m.leaf, m.icoord)


# cells compared: interior and the non-corner ghosts of both layers (the
# four 2x2 corner blocks are never read by the scheme and the device does
# not fill them)
MASK = np.zeros((20, 20), dtype=bool)
MASK[1:19, 1:19] = True
for _a in (0, 1, 18, 19):
    for _b in (0, 1, 18, 19):
        MASK[_a, _b] = False
MASK[0, 2:18] = MASK[19, 2:18] = MASK[2:18, 0] = MASK[2:18, 19] = True


def digest_u(nbmax, nbleafs, lp, leaf, icoord, u):
    """SHA-256 of the tree (nbmax, nbleafs, lp, leaf, icoord of IDs 1..nbmax,
    rows indexed by ID) and of the leaves' conserved variables on the MASK
    cells, -0.0 taken as +0.0."""
    n = int(nbmax)
    lv = [nb for nb in range(1, n + 1) if leaf[nb]]
    h = hashlib.sha256()
    h.update(np.array([n, int(nbleafs)], dtype=np.int64).tobytes())
    for a in (lp[1:n + 1], leaf[1:n + 1], icoord[1:n + 1]):
        h.update(np.ascontiguousarray(a, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(u[lv][:, :, MASK] + 0.0).tobytes())
    return np.frombuffer(h.digest(), dtype=np.uint8).copy()
