"""The oracle against golden vectors written once from the reference's own
hot-path Fortran (tests/golden/make_golden.py over oracle/_ref).  Runs where
the reference tree is absent; inputs are regenerated from fixed seeds."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import inputs as I  # noqa: E402
from wxoracle import Oracle  # noqa: E402

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz"))
O = Oracle()


@pytest.mark.parametrize("variant,choice", I.FLUX_CASES)
def test_fluxes_golden(variant, choice):
    cfg = I.flux_cfg(variant)
    pp = cfg.params()
    m = I.flux_mesh(O, cfg)
    f, g, err = O.fluxes(pp, m, 1, choice)
    assert np.array_equal(f[:, 1:18, 1:18], G[f"flux_{variant}_{choice}_f"])
    assert np.array_equal(g[:, 1:18, 1:18], G[f"flux_{variant}_{choice}_g"])


def test_atomic_aloss_cooldmc_golden():
    pp3 = I.cfg("c3").params()
    states, dts = I.atomic_inputs()
    got = np.stack([O.atomic(pp3, dt, s) for s, dt in zip(states, dts)])
    assert np.array_equal(got, G["atomic"])
    assert np.array_equal(np.array([O.aloss(*a) for a in I.aloss_inputs()]), G["aloss"])
    pp2 = I.cfg("c2").params()
    assert np.array_equal(np.array([O.cooldmc(pp2, T) for T in I.dmc_inputs()]), G["cooldmc"])


@pytest.mark.parametrize("variant", I.LIMITER_VARIANTS)
def test_limiters_golden(variant):
    pp = I.flux_cfg(variant).params()
    res = []
    for a in I.limiter_inputs():
        pl, pr = O.limiter(pp, a[0], a[1], a[2], a[3])
        res.append(np.concatenate([pl, pr]))
    assert np.array_equal(np.array(res), G[f"limiter_{variant}"])


@pytest.mark.parametrize("variant,nsteps", [("c1", 5), ("c2", 4), ("c3", 4)])
def test_steps_golden(variant, nsteps):
    cfg = I.cfg(variant)
    pp, m = I.step_mesh(O, cfg)
    t = 0.0
    dts = []
    for _ in range(nsteps):
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        dts.append(dt)
        if cfg.nlevs > 1:
            O.admesh(pp, m)
    assert np.array_equal(np.array(dts), G[f"step_{variant}_dt"])
    lv = m.leaves()
    if variant == "c1":
        assert np.array_equal(m.u[lv][:, :, 2:18, 2:18], G["step_c1_u"])
    else:
        assert np.array_equal(np.array(lv, dtype=np.int32), G[f"step_{variant}_leaves"])
        np.testing.assert_allclose(m.u[lv][:, :, 2:18, 2:18].sum(axis=(2, 3)), G[f"step_{variant}_sums"],
                                   rtol=1e-13)


# --- mesh layer: vectors from the reference's own locatebounds / admesh /
# time loop (tests/golden/make_golden_mesh.py) ---
import mesh_inputs as MI  # noqa: E402
from harness import oracle_mesh  # noqa: E402

GM = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_mesh_vectors.npz"))


@pytest.mark.parametrize("variant,seed", MI.LB_CASES)
def test_locatebounds_order_golden(variant, seed):
    cfg, pp, m = MI.random_mesh(O, variant, seed)
    n = O.locatebounds(pp, m)
    assert np.array_equal(m.bounds[:n], GM[f"lb_{variant}_{seed}"])


@pytest.mark.parametrize("variant,steps", MI.LOOP_CASES)
def test_time_loop_golden(variant, steps):
    """The oracle's time loop (timestep, tstep, admesh) against the
    reference's: dt and a digest of the tree and leaf state, every step."""
    cfg = MI.cfg_for(variant)
    pp, m = oracle_mesh(O, cfg, 3000)
    if cfg.nlevs > 1:
        for _ in range(3):
            assert O.admesh(pp, m) == 0
    t = 0.0
    dts, dig = GM[f"loop_{variant}_dt"], GM[f"loop_{variant}_digest"]
    assert len(dts) == steps
    for s in range(steps):
        dt = O.timestep(pp, m)
        assert dt == dts[s], (variant, s)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        if cfg.nlevs > 1:
            assert O.admesh(pp, m) == 0
        assert np.array_equal(MI.state_digest(m), dig[s]), (variant, s)
