"""Error behaviour through the C-ABI, as the reference's `print` + `stop`
paths: block capacity exhausted during admesh (src/admesh.f90:44-47,
'nblocks exceeded limit'), a Riemann problem without a valid star speed
(src/hllcfluxes.f90:170-171, 'Error in hllc'), and rejected parameters."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Ensemble, Simulation, WxError

pytestmark = pytest.mark.gpu


def test_capacity_exceeded_during_admesh():
    """The supernova set-up on 8x8 roots refines around the blast on its
    first admesh passes; with room for 90 block IDs that stops with
    WX_EBLOCKS and the reference's message, and the context still closes."""
    cfg = P.config_c2(nblocks=90)
    sim = Simulation(cfg)
    sim.basegrid()
    with pytest.raises(WxError, match=r"status 2\).*nblocks exceeded limit"):
        for _ in range(4):
            sim.admesh()
    sim.close()


def test_nan_state_stops_with_the_hllc_error():
    """A NaN momentum in one cell leaves its faces without a star speed: the
    step returns WX_EHLLC with the reference's 'Error in hllc' text."""
    cfg = P.config_c1()
    sim = Simulation(cfg, nblocks=64)
    sim.basegrid()
    u = sim.download()
    lp, leaf, ic, nbmax, nbleafs = sim.tree()
    u[1, 1, 10, 10] = np.nan
    sim.upload(u, lp, leaf, ic, nbmax, nbleafs)
    with pytest.raises(WxError, match=r"status 1\).*Error in hllc"):
        sim.step(1, admesh=False)
    sim.close()


def test_rejected_parameters():
    """Boundary codes other than outflow/reflective (the reference's INFLOW /
    PERIODIC / BBC options do not work in the reference) are refused by
    wx_create; ensemble instances must share neq; groups lie in [1, n]."""
    import ctypes
    from walicxe2d_amd.capi import last_error, load_library
    L = load_library()
    p = P.config_c1().params()
    p.bc[0] = 2
    h = ctypes.c_void_p()
    assert L.wx_create(ctypes.byref(p), ctypes.byref(h)) == 3
    assert "bc" in last_error()
    with pytest.raises(WxError, match="status 3"):
        Ensemble([P.config_c1(), P.config_c3(nbrootx=4, nbrooty=1, nblocks=100)])
    with pytest.raises(WxError, match="status 3"):
        Ensemble([P.config_c1(), P.config_c1()], groups=3)
