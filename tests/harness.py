"""Shared helpers for the parity tests: build the same initial mesh for the
oracle and the GPU library, and compare states cell by cell."""
from __future__ import annotations

import numpy as np

from walicxe2d_amd import params as P
from wxoracle import MeshState, Oracle

# Cells compared: interior plus the non-corner ghost ring (both layers).
# Corner ghosts depend on the order of the reference's boundary list and are
# never read by the scheme (stencils stay within rows/columns).
MASK = np.zeros((20, 20), dtype=bool)
MASK[1:19, 1:19] = True
for a in (0, 1, 18, 19):
    for b in (0, 1, 18, 19):
        MASK[a, b] = False
MASK_INTERIOR = np.zeros((20, 20), dtype=bool)
MASK_INTERIOR[2:18, 2:18] = True
MASK_RING = MASK.copy()
for a in (0, 19):
    MASK_RING[a, 2:18] = True
    MASK_RING[2:18, a] = True
for a in (0, 1, 18, 19):
    for b in (0, 1, 18, 19):
        MASK_RING[a, b] = False


def oracle_mesh(O: Oracle, cfg: P.Config, cap: int):
    pp = cfg.params()
    m = MeshState(cap, cfg.neq)
    O.basegrid(pp, m)
    u, lp, ic = P.initial_state(cfg)
    n = lp.shape[0] - 1
    m.u[1:n + 1] = u[1:]
    for nb in range(1, n + 1):
        m.primit[nb] = O.calcprim(pp, m.u[nb])
    return pp, m


def cell_scale(b: np.ndarray, gamma: float = 5. / 3.) -> np.ndarray:
    """Per-cell magnitude of each conserved variable, [blocks][neq][20][20]:
    rho for density and scalars, rho*(|v| + c_s) for the momenta (a momentum
    component can cross zero while the cell's momentum scale does not), and
    the total energy for energy."""
    rho = np.maximum(np.abs(b[:, 0]), 1e-300)
    v2 = (b[:, 1] / rho) ** 2 + (b[:, 2] / rho) ** 2     # rho**2 would underflow in unset cells
    eint = np.maximum(b[:, 3] - 0.5 * rho * v2, 1e-300)
    cs = np.sqrt(np.maximum(gamma * (gamma - 1.) * eint / rho, 0.0))
    sc = np.empty_like(b)
    sc[:, 0] = rho
    sc[:, 1] = rho * (np.sqrt(v2) + cs)
    sc[:, 2] = sc[:, 1]
    sc[:, 3] = np.abs(b[:, 3])
    for k in range(4, b.shape[1]):
        sc[:, k] = rho
    return sc


def rel_err(a: np.ndarray, b: np.ndarray, mask=MASK, nvar=None):
    """max over masked cells and the first nvar variables (all if None) of
    |a-b| / scale(cell, var), where the scale is the cell's own magnitude of
    that variable (cell_scale)."""
    sc = cell_scale(b)
    d = np.abs(a - b) / np.maximum(sc, 1e-300)
    if nvar is not None:
        d = d[:, :nvar]
    d = d[:, :, mask]
    if d.size == 0:
        return 0.0, None
    i = int(np.argmax(d))
    idx = np.unravel_index(i, d.shape)
    return float(d.max()), (int(idx[1]), int(idx[0]), int(idx[2]))
