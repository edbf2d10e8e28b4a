"""Legacy-VTK output (walicxe2d_amd/vtkout.py, after src/output.f90:73-256)
of an oracle state on a 3-level cooling-jet mesh: the file parses as the
legacy format, the geometry tiles the box, every cell is an axis-aligned
square of its block's spacing, and each field at a cell's centre equals the
oracle's value there (primitives and T from the oracle's uprim)."""
import math

import numpy as np

from walicxe2d_amd import params as P
from walicxe2d_amd import vtkout
from wxoracle import Oracle

from harness import oracle_mesh


def _read(path, endian=">"):
    f4, i4 = np.dtype(endian + "f4"), np.dtype(endian + "i4")
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    def arr(dt, n):
        nonlocal pos
        a = np.frombuffer(data, dtype=dt, count=n, offset=pos)
        pos += a.nbytes
        assert data[pos:pos + 1] == b"\n"
        pos += 1
        return a

    out = {"head": [line() for _ in range(4)]}
    tok = line().split()
    assert tok[0] == "POINTS" and tok[2] == "float"
    out["points"] = arr(f4, 3 * int(tok[1])).reshape(-1, 3)
    tok = line().split()
    assert tok[0] == "CELLS"
    nc = int(tok[1])
    out["cells"] = arr(i4, int(tok[2])).reshape(nc, 5)
    tok = line().split()
    assert tok == ["CELL_TYPES", str(nc)]
    out["types"] = arr(i4, nc)
    assert line().split() == ["CELL_DATA", str(nc)]
    fields = {}
    while pos < len(data):
        tok = line().split()
        if tok[0] == "SCALARS":
            assert tok[2:] == ["float", "1"] and line() == "LOOKUP_TABLE default"
            fields[tok[1]] = arr(f4, nc)
        else:
            assert tok == ["VECTORS", "vel", "float"]
            fields["vel"] = arr(f4, 3 * nc).reshape(nc, 3)
    out["fields"] = fields
    return out


def _state():
    cfg = P.config_c3(nbrootx=4, nbrooty=2, nblocks=3000)
    cfg.nlevs = 3
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 3000)
    t = 0.0
    for _ in range(40):                                   # the jet enters and the mesh refines
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        assert O.admesh(pp, m) == 0
    return cfg, O, pp, m


def test_vtk_geometry_and_fields(tmp_path):
    cfg, O, pp, m = _state()
    nb_max = m.nbmax
    fname = str(tmp_path / "out000.vtk")
    nc = vtkout.write_vtk(fname, cfg, m.u, m.lp, m.leaf, m.icoord, nb_max)
    ids = [nb for nb in range(1, nb_max + 1) if m.leaf[nb]]
    levels = sorted({int(m.lp[nb, 0]) for nb in ids})
    assert len(levels) >= 2, levels                      # a refined mesh
    assert nc == 256 * len(ids)
    v = _read(fname)
    assert v["head"] == ["# vtk DataFile Version 4.2 ", "output from walicxe", "BINARY",
                         "DATASET UNSTRUCTURED_GRID"]
    pts, cells = v["points"].astype(np.float64), v["cells"]
    assert len(pts) == 289 * len(ids)
    assert np.all(v["types"] == 8) and np.all(cells[:, 0] == 4)
    sc = cfg.sc
    rsc = sc["rsc"]
    dx, dy = cfg.dxdy()
    # every cell: an axis-aligned square whose side is its block's spacing,
    # and the cells tile the box exactly once
    c = pts[cells[:, 1:]]                                 # [nc][4][3]
    x0, x1 = c[:, :, 0].min(1), c[:, :, 0].max(1)
    y0, y1 = c[:, :, 1].min(1), c[:, :, 1].max(1)
    blk = np.repeat(np.asarray(ids), 256)
    lev = m.lp[blk, 0]
    hx = np.asarray(dx)[lev - 1] * rsc
    hy = np.asarray(dy)[lev - 1] * rsc
    assert np.allclose(x1 - x0, hx, rtol=1e-5) and np.allclose(y1 - y0, hy, rtol=1e-5)   # float32 corners
    assert math.isclose(((x1 - x0) * (y1 - y0)).sum(), cfg.xmax * cfg.ymax * rsc * rsc, rel_tol=1e-5)
    # each cell's indices (i, j) in its block from its lower-left corner
    i = np.rint(x0 / hx).astype(int) - m.icoord[blk, 0] + 1
    j = np.rint(y0 / hy).astype(int) - m.icoord[blk, 1] + 1
    assert i.min() == 1 and i.max() == 16 and j.min() == 1 and j.max() == 16
    f = v["fields"]
    assert list(f) == ["rho", "P", "rho_HI", "T", "blockID", "vel"]
    assert np.array_equal(f["blockID"], blk.astype(np.float32))
    u = m.u[blk, :, j + 1, i + 1]                         # [nc][neq]; tile [id][var][j][i], cell 1 at index 2
    assert np.array_equal(f["rho"], (u[:, 0] * sc["rhosc"]).astype(np.float32))
    rng = np.random.default_rng(3)
    vs = math.sqrt(sc["vsc2"])
    for k in rng.choice(nc, 300, replace=False):
        prim, T = O.uprim(pp, u[k])
        assert math.isclose(f["T"][k], T, rel_tol=1e-6), (k, f["T"][k], T)
        assert math.isclose(f["P"][k], prim[3], rel_tol=1e-6)
        assert math.isclose(f["rho_HI"][k], prim[4], rel_tol=1e-6)
        assert np.allclose(f["vel"][k], [prim[1] * vs, prim[2] * vs, 0.0], rtol=1e-6, atol=1e-30)


def test_vtk_little_endian_is_the_byte_swap(tmp_path):
    cfg, O, pp, m = _state()
    a, b = str(tmp_path / "be.vtk"), str(tmp_path / "le.vtk")
    vtkout.write_vtk(a, cfg, m.u, m.lp, m.leaf, m.icoord, m.nbmax)
    vtkout.write_vtk(b, cfg, m.u, m.lp, m.leaf, m.icoord, m.nbmax, endian="<")
    va, vb = _read(a), _read(b, "<")
    assert np.array_equal(va["points"], vb["points"]) and np.array_equal(va["cells"], vb["cells"])
    for k in va["fields"]:
        assert np.array_equal(va["fields"][k], vb["fields"][k])
    assert open(a, "rb").read() != open(b, "rb").read()
