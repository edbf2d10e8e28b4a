"""Legacy-VTK output of a mesh state, after the reference's OUTVTK branch
(src/output.f90:73-256): one unstructured grid of all leaf blocks (block-ID
order), (nx+1)(ny+1) points per block (i outer, j inner), one pixel cell per
interior cell, and the cell fields rho, P, rho_HI, T, blockID, Passive and
the velocity vector.

Kept from the reference: the header lines, the point, connectivity and
cell-type records and their order (`k, k+1, k+ny+1, k+ny+2`, type 8), the
fields, their names and order, and their scalings (rho * rhosc, P and the
scalar in code units, T from the primitive recovery of the run's cooling
mode, velocities * sqrt(vsc2), coordinates * rsc).  Differences, each a
choice of this writer:
  * numbers are big-endian 32-bit as the legacy format prescribes (the
    reference opens the file LITTLE_ENDIAN and writes default-kind reals,
    8 bytes in its -r8 builds; `endian="<"` gives the little-endian bytes);
  * blockID is written as a float, as its header declares (the reference
    writes the integer's bytes);
  * rho_HI and Passive are written only when the state has those
    variables (neq >= 5 and >= 6);
  * counts in the header lines are separated by one space (the reference's
    `i` edit descriptor pads them).
"""
from __future__ import annotations

import math

import numpy as np

from .params import COOLING, NX, NY


def primitives(cfg, u: np.ndarray):
    """Primitives and temperature of conserved tiles [..., neq, j, i], the
    device's primitive recovery (src/uprim.f90:5-63) in numpy, term by term."""
    sc = cfg.sc
    cv, tempsc = sc["cv"], sc["Tempsc"]
    neq = u.shape[-3]
    q = np.empty_like(u)
    r = np.maximum(u[..., 0, :, :], 1e-15)
    q[..., 0, :, :] = r
    q[..., 1, :, :] = u[..., 1, :, :] / r
    q[..., 2, :, :] = u[..., 2, :, :] / r
    for k in range(4, neq):
        q[..., k, :, :] = u[..., k, :, :]
    p = (u[..., 3, :, :] - 0.5 * r * (q[..., 1, :, :] * q[..., 1, :, :] + q[..., 2, :, :] * q[..., 2, :, :])) / cv
    p = np.maximum(p, 1e-15)
    mode = COOLING[cfg.cooling]
    if mode == 0:
        T = p / r
    elif mode == 1:
        dentot = np.maximum(2.0 * r - (q[..., 4, :, :] if neq > 4 else 0.0), 1e-15)
        T = np.maximum(1.0, (p / dentot) * tempsc)
        p = dentot * T / tempsc
    else:
        T = np.maximum(1.0, (p / r) * tempsc)
        p = r * T / tempsc
    q[..., 3, :, :] = p
    return q, T


def write_vtk(fname: str, cfg, u: np.ndarray, lp: np.ndarray, leaf: np.ndarray, icoord: np.ndarray,
              nbmax: int | None = None, endian: str = ">") -> int:
    """Writes the leaves of a state to fname; returns the number of cells.

    u: conserved tiles [nbmax+1][neq][20][20] ([id][var][j][i], slot = block
    ID, interior cell (i, j) at [j+1][i+1]); lp [nbmax+1][8] (lp(nb, k) at lp[nb, k-1]), leaf [nbmax+1],
    icoord [nbmax+1][2], as Simulation.download / tree return them."""
    if nbmax is None:
        nbmax = len(leaf) - 1
    ids = [nb for nb in range(1, nbmax + 1) if leaf[nb]]
    sc = cfg.sc
    rsc, rhosc, vsc = sc["rsc"], sc["rhosc"], math.sqrt(sc["vsc2"])
    dx, dy = cfg.dxdy()
    neq = u.shape[1]
    f4, i4 = np.dtype(endian + "f4"), np.dtype(endian + "i4")
    nl = len(ids)
    npoints = nl * (NX + 1) * (NY + 1)
    ncells = nl * NX * NY
    ii = np.arange(NX + 1)[:, None]
    jj = np.arange(NY + 1)[None, :]
    with open(fname, "wb") as fh:
        w = fh.write
        w(b"# vtk DataFile Version 4.2 \n")
        w(b"output from walicxe\n")
        w(b"BINARY\n")
        w(b"DATASET UNSTRUCTURED_GRID\n")
        w(f"POINTS {npoints} float\n".encode())
        for nb in ids:                                   # src/output.f90:117-125
            lev = int(lp[nb, 0])
            pts = np.zeros((NX + 1, NY + 1, 3))
            pts[..., 0] = (icoord[nb, 0] + ii).astype(np.float64) * dx[lev - 1] * rsc
            pts[..., 1] = (icoord[nb, 1] + jj).astype(np.float64) * dy[lev - 1] * rsc
            w(pts.astype(f4).tobytes())
        w(b"\n")
        w(f"CELLS {ncells} {ncells * 5}\n".encode())
        k = 0                                            # src/output.f90:130-138
        cell = []
        for _ in ids:
            for i in range(NX + 1):
                for j in range(NY + 1):
                    if i < NX and j < NY:
                        cell.append((4, k, k + 1, k + NY + 1, k + NY + 2))
                    k += 1
        w(np.asarray(cell, dtype=np.int64).astype(i4).tobytes())
        w(b"\n")
        w(f"CELL_TYPES {ncells}\n".encode())
        w(np.full(ncells, 8, dtype=i4).tobytes())
        w(b"\n")
        w(f"CELL_DATA {ncells}\n".encode())
        tiles = u[ids][:, :, 2:2 + NY, 2:2 + NX]         # interior [n][neq][j][i], 2 ghost cells a side
        q, T = primitives(cfg, tiles)

        def field(name, a):                              # a: [n][j][i] -> i outer, j inner
            w(f"SCALARS {name} float 1\n".encode())
            w(b"LOOKUP_TABLE default\n")
            w(np.ascontiguousarray(np.swapaxes(a, 1, 2)).astype(f4).tobytes())
            w(b"\n")
        field("rho", tiles[:, 0] * rhosc)
        field("P", q[:, 3])
        if neq > 4:
            field("rho_HI", q[:, 4])
        field("T", T)
        field("blockID", np.broadcast_to(np.asarray(ids, dtype=np.float64)[:, None, None], T.shape))
        if neq > 5:
            field("Passive", q[:, 5])
        w(b"VECTORS vel float\n")
        vel = np.stack([np.swapaxes(q[:, 1] * vsc, 1, 2), np.swapaxes(q[:, 2] * vsc, 1, 2),
                        np.zeros((nl, NX, NY))], axis=-1)
        w(vel.astype(f4).tobytes())
        w(b"\n")
    return ncells


def save(sim, fname: str, endian: str = ">") -> int:
    """Writes a Simulation's current state (Simulation.download / tree)."""
    u = sim.download()
    lp, leaf, ic, nbmax, _ = sim.tree()
    return write_vtk(fname, sim.cfg, u, lp, leaf, ic, nbmax, endian)
