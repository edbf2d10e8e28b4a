"""The reference's BIN checkpoint files (OUTBIN, src/output.f90:39-67) and
their Python readers (py/walicxe2d_utils.py, Python 2 and per-value
struct.unpack there), here for Python 3 with numpy.

File names follow the reference: with MPI (`src/output.f90:44-45`)
`<path>/BIN/pointers<rank:03>.<itprint:03>.bin` and `blocks...`, without
MPI `pointers<itprint:03>.bin`.  The files themselves are written and read
by the library (`wx_write_bin` / `wx_read_bin`, and `wx_ens_write_bin` for
one instance of an ensemble): Simulation.write_bin / read_bin,
Ensemble.write_bin, or `save` / `warm_start` below.
"""
from __future__ import annotations

import os

import numpy as np


def names(path: str, itprint: int, rank: int | None = 0):
    """(pointers, blocks) file names of one output; rank None: serial names."""
    b = os.path.join(path, "BIN")
    if rank is None:
        return (os.path.join(b, f"pointers{itprint:03d}.bin"), os.path.join(b, f"blocks{itprint:03d}.bin"))
    return (os.path.join(b, f"pointers{rank:03d}.{itprint:03d}.bin"),
            os.path.join(b, f"blocks{rank:03d}.{itprint:03d}.bin"))


def save(sim, path: str, itprint: int, rank: int | None = 0, logical_true: int = -1):
    """Writes output number itprint of a Simulation as the reference does."""
    p, b = names(path, itprint, rank)
    os.makedirs(os.path.dirname(p), exist_ok=True)
    sim.write_bin(p, b, logical_true)
    return p, b


def warm_start(sim, path: str, itprint: int, nproc: int = 1, serial: bool = False):
    """src/basegrid.f90:167-251: replaces the Simulation's mesh and state by
    output itprint of an nproc-rank run (one pair per rank)."""
    pairs = [names(path, itprint, None if serial else r) for r in range(nproc)]
    sim.read_bin([a for a, _ in pairs], [b for _, b in pairs])


def read_pointers(fname: str):
    """(nbmin, nbmax, nbleafs, nblocks, lp[n,8], leaf[n] as 0/1, icoord[n,2])
    of one pointers file, n = nbmax - nbmin + 1."""
    raw = np.fromfile(fname, dtype="<i4")
    nbmin, nbmax, nbleafs, nblocks = (int(x) for x in raw[:4])
    n = nbmax - nbmin + 1
    lp = raw[4:4 + 8 * n].reshape(8, n).T
    lg = raw[4 + 8 * n:4 + 9 * n]
    ic = raw[4 + 9 * n:4 + 11 * n].reshape(2, n).T
    return nbmin, nbmax, nbleafs, nblocks, lp, (lg != 0).astype(np.int32), ic


def read_mesh(nproc, nx, ny, nlevs, nout, xmax, ymax, nbrootx, nbrooty, path="./"):
    """py/walicxe2d_utils.py read_mesh: one row per leaf, (x0, y0, x1, y1,
    level, rank).  The original takes a leaf as -1 only (ifort's .true.);
    any nonzero value counts here, so flang's 1 reads as well."""
    dx = xmax / nbrootx / nx / 2.0 ** np.arange(nlevs)
    dy = ymax / nbrooty / ny / 2.0 ** np.arange(nlevs)
    rows = []
    for rank in range(nproc):
        _, _, _, _, lp, leaf, ic = read_pointers(os.path.join(path, f"pointers{rank:03d}.{nout:03d}.bin"))
        for q in np.nonzero(leaf)[0]:
            lev = lp[q, 0] - 1
            rows.append((ic[q, 0] * dx[lev], ic[q, 1] * dy[lev], (ic[q, 0] + nx) * dx[lev],
                         (ic[q, 1] + ny) * dy[lev], lp[q, 0], rank))
    return np.array(rows, dtype=np.float64).reshape(-1, 6)


def read_blocks(nproc, nx, ny, nlevs, neqs, nout, xmax, ymax, nbrootx, nbrooty, path="./", double=True):
    """py/walicxe2d_utils.py read_blocks: array [nx+4, ny+4, neqs, leaves]
    indexed [i, j, eq, leaf] over the padded block (index 0 = cell -1)."""
    out = []
    for rank in range(nproc):
        _, _, nbleafs, _, _, _, _ = read_pointers(os.path.join(path, f"pointers{rank:03d}.{nout:03d}.bin"))
        raw = np.fromfile(os.path.join(path, f"blocks{rank:03d}.{nout:03d}.bin"), dtype="<f8" if double else "<f4")
        raw = raw[:nbleafs * (ny + 4) * (nx + 4) * neqs].reshape(nbleafs, ny + 4, nx + 4, neqs)   # [leaf][j][i][eq]
        out.append(raw.transpose(2, 1, 3, 0))
    return np.concatenate(out, axis=3) if out else np.zeros((nx + 4, ny + 4, neqs, 0))


def read_map(nproc, nx, ny, nlevs, neqs, neq, nout, xmax, ymax, nbrootx, nbrooty, path="./", double=True):
    """py/walicxe2d_utils.py read_map: variable neq on the finest uniform
    grid, returned [y][x] as the original's map.T; finest-level leaves are
    copied, coarser ones zoomed with scipy.ndimage.zoom(order=1) as there."""
    import scipy.ndimage
    nxtot = nbrootx * nx * 2 ** (nlevs - 1)
    nytot = nbrooty * ny * 2 ** (nlevs - 1)
    dx, dy = xmax / nxtot, ymax / nytot
    out = np.zeros((nxtot, nytot), dtype=np.float64 if double else np.float32)
    mesh = read_mesh(nproc, nx, ny, nlevs, nout, xmax, ymax, nbrootx, nbrooty, path=path)
    blocks = read_blocks(nproc, nx, ny, nlevs, neqs, nout, xmax, ymax, nbrootx, nbrooty, path=path, double=double)
    for nb in range(mesh.shape[0]):
        x0 = int(mesh[nb, 0] / dx)
        y0 = int(mesh[nb, 1] / dy)
        a = blocks[2:ny + 2, 2:nx + 2, neq, nb]
        if int(mesh[nb, 4]) == nlevs:
            out[x0:x0 + nx, y0:y0 + ny] = a
        else:
            f = 2 ** (nlevs - int(mesh[nb, 4]))
            out[x0:x0 + f * nx, y0:y0 + f * ny] = scipy.ndimage.zoom(a, f, order=1)
    return out.T
