"""Warm start on the GPU: a jet run checkpointed with wx_write_bin after 30
steps, read into a fresh context (wx_read_bin, src/basegrid.f90:167-251)
and continued for 15 steps, equals the uninterrupted run bit for bit."""
import numpy as np
import pytest

from walicxe2d_amd import binio
from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Simulation

from harness import MASK

pytestmark = pytest.mark.gpu


def test_warm_start_continues_the_run(tmp_path):
    cfg = P.config_c3(nbrootx=16, nbrooty=4, nblocks=4000)
    a = Simulation(cfg)
    a.basegrid()
    for _ in range(4):
        a.admesh()
    a.step(30)
    binio.save(a, str(tmp_path), 1)
    b = Simulation(cfg)
    binio.warm_start(b, str(tmp_path), 1)
    b.set_time(a.time)
    lpa, leafa, ica, nba, nla = a.tree()
    lpb, leafb, icb, nbb, nlb = b.tree()
    assert (nba, nla) == (nbb, nlb) and np.array_equal(lpa, lpb) and np.array_equal(leafa, leafb)
    ids = np.nonzero(leafa)[0]
    assert np.array_equal(a.download()[ids], b.download()[ids])
    for s in range(15):
        ta, dta, _ = a.step(1)
        tb, dtb, _ = b.step(1)
        assert (ta, dta) == (tb, dtb), s
        lpa, leafa, _, nba, _ = a.tree()
        lpb, leafb, _, nbb, _ = b.tree()
        assert nba == nbb and np.array_equal(lpa, lpb) and np.array_equal(leafa, leafb), s
        ids = np.nonzero(leafa)[0]
        ua, ub = a.download()[ids], b.download()[ids]
        assert np.array_equal(ua[:, :, MASK], ub[:, :, MASK]), s
    a.close()
    b.close()


def test_ensemble_instance_checkpoint_equals_single_run(tmp_path):
    """wx_ens_write_bin of an instance writes the same bytes as wx_write_bin
    of the same mesh run alone (the ensemble is bit-identical to it)."""
    from walicxe2d_amd.capi import Ensemble
    cfgs = [P.config_c4(i, nblocks=3000) for i in (0, 9)]
    e = Ensemble(cfgs, groups=2)
    e.basegrid()
    for _ in range(3):
        e.admesh()
    e.step(5)
    s = Simulation(cfgs[1])
    s.basegrid()
    for _ in range(3):
        s.admesh()
    s.step(5)
    e.write_bin(1, str(tmp_path / "pe.bin"), str(tmp_path / "be.bin"))
    s.write_bin(str(tmp_path / "ps.bin"), str(tmp_path / "bs.bin"))
    assert (tmp_path / "pe.bin").read_bytes() == (tmp_path / "ps.bin").read_bytes()
    assert (tmp_path / "be.bin").read_bytes() == (tmp_path / "bs.bin").read_bytes()
    e.close()
    s.close()
