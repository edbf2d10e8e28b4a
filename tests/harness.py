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


def rel_err(a: np.ndarray, b: np.ndarray, mask=MASK, floor_rel=1e-8):
    """max over masked cells of |a-b| / max(|b|, floor_rel * max|b_var|), per variable."""
    worst = 0.0
    where = None
    for k in range(a.shape[1]):
        ak = a[:, k][:, mask]
        bk = b[:, k][:, mask]
        scale = max(np.max(np.abs(bk)), 1e-300)
        d = np.abs(ak - bk) / np.maximum(np.abs(bk), floor_rel * scale)
        if d.size and d.max() > worst:
            worst = float(d.max())
            where = (k, np.unravel_index(int(np.argmax(d)), d.shape))
    return worst, where
