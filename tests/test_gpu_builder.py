"""The device table builder (walicxe2d_amd/csrc/wx_lbuild.hip) against the
host compilation of the same list (wx_set_builder 0: locatebounds +
compile_ops, src/locatebounds.f90:29-540): identical leaf records and ghost
ops of both fills after every mesh change of an AMR run, and bit-identical
states at the end.  Workloads: the cooling jet (jet imposition flags), the
point explosion in the OUTFYB compat mode of src/boundaryI.f90:87 (fill I's
own table), the 5-level DMC case, the many-workgroup scan, and the default
choice by mesh size across its threshold."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Simulation

pytestmark = pytest.mark.gpu


def _pair(cfg, steps, cap, check_every, mode=1):
    sims = []
    for on_device in (mode, 0):
        s = Simulation(cfg, nblocks=cap)
        s.set_builder(on_device)
        s.basegrid()
        for _ in range(4):
            s.admesh()
        sims.append(s)
    dev, host = sims
    checked = 0
    for k in range(0, steps, check_every):
        n = min(check_every, steps - k)
        dev.step(n, admesh=True)
        host.step(n, admesh=True)
        td, th = dev.tables(), host.tables()
        assert td[0].shape == th[0].shape, (k, td[0].shape, th[0].shape)
        for a, b, what in zip(td, th, ("leaf records", "fill II ops", "fill I ops")):
            assert np.array_equal(a, b), f"step {k + n}: {what} differ at {np.argwhere(a != b)[:4].tolist()}"
        assert dev.info() == host.info(), (k, dev.info(), host.info())
        checked += 1
    ud, uh = dev.download(), host.download()
    assert dev.tree()[3] == host.tree()[3]
    assert np.array_equal(ud.view(np.int64), uh.view(np.int64)), "states differ"
    nleaf = td[0].shape[0]
    for s in sims:
        s.close()
    print(f"{cfg.name}: {checked} table comparisons, {nleaf} leaves at the end, states bit-identical")
    return nleaf


def test_builder_jet_c3():
    cfg = P.config_c3()
    assert _pair(cfg, 300, cfg.nblocks, 25) > 64


def test_builder_outfyb_c1():
    # the point explosion with 3 levels of refinement, so the OUTFYB tables
    # are rebuilt as the blast moves
    cfg = P.config_c1(nlevs=3, nblocks=4000)
    assert cfg.params().compat & 2
    _pair(cfg, 150, cfg.nblocks, 10)


def test_builder_dmc_c2():
    cfg = P.config_c2(nblocks=20000)
    _pair(cfg, 120, cfg.nblocks, 20)


def test_builder_many_workgroup_scan():
    # the form the scan takes above 16384 block IDs, forced on the jet
    cfg = P.config_c3()
    _pair(cfg, 200, cfg.nblocks, 50, mode=2)


def test_builder_by_size_crossing():
    # the default mode (3) on a point explosion whose mesh grows from 3012
    # leaves (54 x 54 roots after the initial refinement) past 3072 near
    # step 30: the host list first, then the device builder, against the
    # host list throughout
    cfg = P.config_c1(nbrootx=54, nbrooty=54, nlevs=3, nblocks=30000)
    assert _pair(cfg, 60, cfg.nblocks, 5, mode=3) >= 3072


def test_builder_list_length():
    cfg = P.config_c3()
    n = []
    for on_device in (1, 2, 0):
        s = Simulation(cfg, nblocks=cfg.nblocks)
        s.set_builder(on_device)
        s.basegrid()
        for _ in range(4):
            s.admesh()
        n.append(s.locatebounds())
        s.close()
    assert n[0] == n[1] == n[2] > 0, n
