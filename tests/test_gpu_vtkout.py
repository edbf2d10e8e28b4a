"""Legacy-VTK output of a device run (vtkout.save through the C-ABI's
download and tree) against the same writer on the oracle's state after the
same steps: identical geometry, connectivity and block IDs, and every field
equal to float32 rounding."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd import vtkout
from walicxe2d_amd.capi import Simulation
from wxoracle import Oracle

from harness import oracle_mesh
from test_vtkout import _read

pytestmark = pytest.mark.gpu


def test_vtk_of_device_run_matches_oracle(tmp_path):
    cfg = P.config_c3(nbrootx=4, nbrooty=2, nblocks=3000)
    cfg.nlevs = 3
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 3000)
    sim = Simulation(cfg, nblocks=3000)
    sim.basegrid()
    t = 0.0
    for _ in range(40):                                   # the jet enters and the mesh refines
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        sim.step(1, admesh=False)
        assert O.admesh(pp, m) == 0
        sim.admesh()
    a, b = str(tmp_path / "gpu.vtk"), str(tmp_path / "oracle.vtk")
    nc = vtkout.save(sim, a)
    sim.close()
    assert vtkout.write_vtk(b, cfg, m.u, m.lp, m.leaf, m.icoord, m.nbmax) == nc
    va, vb = _read(a), _read(b)
    assert len({int(m.lp[nb, 0]) for nb in range(1, m.nbmax + 1) if m.leaf[nb]}) >= 2
    assert va["head"] == vb["head"]
    assert np.array_equal(va["points"], vb["points"]) and np.array_equal(va["cells"], vb["cells"])
    assert list(va["fields"]) == list(vb["fields"])
    assert np.array_equal(va["fields"]["blockID"], vb["fields"]["blockID"])
    # the states agree to 1e-11 of each cell's scale (density for n_HI, the
    # flow speed for the velocity), and float32 rounding of two such values
    # may differ by an ulp (2^-23 relative)
    rho = vb["fields"]["rho"].astype(np.float64) / cfg.sc["rhosc"]
    vmax = np.abs(vb["fields"]["vel"]).max()
    atol = {"rho": 0.0, "P": 0.0, "T": 0.0, "rho_HI": 1e-10 * rho, "vel": 1e-10 * vmax}
    for k in ("rho", "P", "rho_HI", "T", "vel"):
        x, y = va["fields"][k].astype(np.float64), vb["fields"][k].astype(np.float64)
        lim = 2.5e-7 * np.abs(y) + atol[k]
        assert np.all(np.abs(x - y) <= lim), (k, np.abs(x - y).max())
