"""Relative L1 density error between two AMR states after projecting both to
the finest uniform grid by piecewise-constant injection:
    L1 = sum |rho - rho_ref| / sum |rho_ref|.
A leaf of level L with lower-left cell index (icx, icy) at its own level
(src/refine.f90:227-261 convention) covers finest cells
[icx f, (icx + 16) f) x [icy f, (icy + 16) f) with f = 2^(nlevs - L).
"""
from __future__ import annotations

import numpy as np

from .params import NX, NY, Config


def project(cfg: Config, u: np.ndarray, lp: np.ndarray, leaf: np.ndarray, icoord: np.ndarray,
            var: int = 0) -> np.ndarray:
    """Variable `var` of every leaf interior on the finest grid [y][x]."""
    top = 2 ** (cfg.nlevs - 1)
    grid = np.full((cfg.nbrooty * NY * top, cfg.nbrootx * NX * top), np.nan)
    for nb in np.nonzero(leaf)[0]:
        lev = int(lp[nb, 0])
        f = 2 ** (cfg.nlevs - lev)
        i0, j0 = int(icoord[nb, 0]) * f, int(icoord[nb, 1]) * f
        blk = u[nb, var, 2:2 + NY, 2:2 + NX]
        grid[j0:j0 + NY * f, i0:i0 + NX * f] = np.repeat(np.repeat(blk, f, axis=0), f, axis=1)
    if np.isnan(grid).any():
        raise ValueError("the leaves do not tile the domain")
    return grid


def l1_density(cfg: Config, state, ref) -> float:
    """state, ref: (u, lp, leaf, icoord) as Simulation.download() / tree() give them."""
    a = project(cfg, *state)
    b = project(cfg, *ref)
    return float(np.abs(a - b).sum() / np.abs(b).sum())
