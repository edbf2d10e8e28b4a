"""The L1 projection used for full-run density errors: closed forms on small
hand-built meshes (no GPU)."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.l1 import l1_density, project


def _two_level_mesh():
    """2x1 roots; root 1 refined into 4 sons (IDs 3..6), root 2 a leaf."""
    cfg = P.config_c1(nbrootx=2, nbrooty=1, nlevs=2)
    lp = np.zeros((7, 8), np.int32)
    leaf = np.zeros(7, np.int32)
    ic = np.zeros((7, 2), np.int32)
    lp[1, 0] = lp[2, 0] = 1
    ic[2] = (16, 0)
    leaf[2] = 1
    for s, (dx, dy) in zip(range(3, 7), ((0, 0), (16, 0), (0, 16), (16, 16))):
        lp[s, 0] = 2
        ic[s] = (dx, dy)
        leaf[s] = 1
    u = np.zeros((7, 4, 20, 20))
    for nb in range(1, 7):
        u[nb, 0] = nb
    return cfg, u, lp, leaf, ic


def test_projection_tiles_and_injects():
    cfg, u, lp, leaf, ic = _two_level_mesh()
    g = project(cfg, u, lp, leaf, ic)
    assert g.shape == (32, 64)
    assert (g[:16, :16] == 3).all() and (g[:16, 16:32] == 4).all()
    assert (g[16:, :16] == 5).all() and (g[16:, 16:32] == 6).all()
    assert (g[:, 32:] == 2).all()   # the coarse leaf covers 32x32 finest cells


def test_l1_closed_form():
    cfg, u, lp, leaf, ic = _two_level_mesh()
    v = u.copy()
    v[2, 0] += 0.5                   # coarse leaf: 1024 finest cells off by 0.5
    ref = (u, lp, leaf, ic)
    err = l1_density(cfg, (v, lp, leaf, ic), ref)
    total = 256 * (3 + 4 + 5 + 6) + 1024 * 2
    assert err == pytest.approx(1024 * 0.5 / total, rel=1e-14)
    assert l1_density(cfg, ref, ref) == 0.0


def test_gaps_are_rejected():
    cfg, u, lp, leaf, ic = _two_level_mesh()
    leaf[4] = 0
    with pytest.raises(ValueError):
        project(cfg, u, lp, leaf, ic)
