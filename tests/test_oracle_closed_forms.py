"""The oracle against closed forms, invariants and independent solutions.

These pin the oracle without the reference binary (see also
tests/test_oracle_golden.py for committed vectors and tests/test_oracle_ref.py
for the reference build itself)."""
import math

import numpy as np
import pytest

from walicxe2d_amd import params as P
from wxoracle import MeshState, Oracle

from harness import oracle_mesh

O = Oracle()


def cfg_ad(**kw):
    base = dict(name="t", neq=5, nlevs=3, nbrootx=4, nbrooty=4, xmax=1.0, ymax=1.0, xphys=P.PC,
                amol=0.6, cooling="ADIABATIC", bc=("REFLECTIVE",) * 4, ic="explosion_code",
                ic_args=dict(rho=1.0, p=1e-2, E=1.0, r=0.1, xc=0.4, yc=0.45), nblocks=2000)
    base.update(kw)
    return P.Config(**base)


def prims(rng, n, neq=5, vmax=3.0):
    q = np.empty((n, neq))
    q[:, 0] = 10 ** rng.uniform(-2, 2, n)
    q[:, 1:3] = rng.uniform(-vmax, vmax, (n, 2))
    q[:, 3] = 10 ** rng.uniform(-3, 2, n)
    q[:, 4:] = q[:, :1] * rng.uniform(0, 1, (n, neq - 4))
    return q


def mirror(q):
    m = q.copy()
    m[..., 1] = -m[..., 1]
    return m


@pytest.mark.parametrize("solver", ["HLLC", "HLL", "HLL_HLLC"])
def test_consistency_F_of_U_U(solver):
    """Any consistent solver returns the physical flux for identical states."""
    pp = cfg_ad(solver=solver).params()
    rng = np.random.default_rng(0)
    for q in prims(rng, 400):
        ff, err = O.riemann(pp, q, q)
        assert err == 0
        np.testing.assert_allclose(ff, O.primf(pp, q), rtol=1e-12, atol=1e-12 * np.abs(O.primf(pp, q)).max())


@pytest.mark.parametrize("solver", ["HLLC", "HLL", "HLL_HLLC"])
def test_supersonic_upwinding(solver):
    """sl > 0 returns F(U_L) and sr < 0 returns F(U_R) exactly."""
    pp = cfg_ad(solver=solver).params()
    rng = np.random.default_rng(1)
    for q in prims(rng, 200):
        l, r = q.copy(), q.copy()
        r[0] *= 2.0
        cs = math.sqrt(pp.gamma * max(l[3], r[3]) / min(l[0], r[0]))
        l[1] = r[1] = 3 * cs
        assert np.array_equal(O.riemann(pp, l, r)[0], O.primf(pp, l))
        l[1] = r[1] = -3 * cs
        assert np.array_equal(O.riemann(pp, l, r)[0], O.primf(pp, r))


def test_hllc_resolves_contacts_exactly():
    """A contact (equal u and p, different rho) is resolved by HLLC: the flux
    is the physical flux of the upwind state (Toro, ch. 10)."""
    pp = cfg_ad(solver="HLLC", compat=0).params()
    rng = np.random.default_rng(2)
    for _ in range(200):
        u = rng.uniform(-0.5, 0.5)
        p = 10 ** rng.uniform(-1, 1)
        l = np.array([10 ** rng.uniform(-1, 1), u, rng.uniform(-1, 1), p, 0.3])
        r = np.array([10 ** rng.uniform(-1, 1), u, rng.uniform(-1, 1), p, 0.7])
        ff, _ = O.riemann(pp, l, r)
        up = l if u >= 0 else r
        np.testing.assert_allclose(ff, O.primf(pp, up), rtol=1e-9, atol=1e-12)


def test_mirror_symmetry_and_typo_location():
    """Mirror symmetry F(m(R), m(L)) = -m(F(L, R)) holds for HLLC with the
    right-star energy typo switched off; with it on (the reference,
    src/hllcfluxes.f90:157) only the energy flux of the right-star branch
    breaks it, and mass / momentum fluxes stay symmetric."""
    rng = np.random.default_rng(3)
    pfix = cfg_ad(solver="HLLC", compat=0).params()
    pref = cfg_ad(solver="HLLC", compat=3).params()
    broke_energy = 0
    for _ in range(2000):
        l, r = prims(rng, 2, vmax=0.5)
        f1, _ = O.riemann(pfix, l, r)
        f2, _ = O.riemann(pfix, mirror(r), mirror(l))
        m = -mirror(f1[None])[0]
        np.testing.assert_allclose(f2, m, rtol=1e-9, atol=1e-12 * (np.abs(f1).max() + 1))
        g1, _ = O.riemann(pref, l, r)
        g2, _ = O.riemann(pref, mirror(r), mirror(l))
        gm = -mirror(g1[None])[0]
        np.testing.assert_allclose(g2[:3], gm[:3], rtol=1e-9, atol=1e-12 * (np.abs(g1).max() + 1))
        if not np.allclose(g2[3], gm[3], rtol=1e-9, atol=1e-12 * (np.abs(g1).max() + 1)):
            broke_energy += 1
    assert broke_energy > 0


@pytest.mark.parametrize("lim", [1, 2, 3, 4, 5, 6])
def test_limiter_identities(lim):
    rng = np.random.default_rng(4)
    for a in rng.normal(0, 1, 300):
        # van Albada: avg(a,a) = a (2a^2+2d)/(2a^2+d) with d = 1e-7; the others give a
        want = a * (2 * a * a + 2e-7) / (2 * a * a + 1e-7) if lim == 3 else a
        assert O.L.wxo_average(lim, a, a) == pytest.approx(want, rel=1e-14)
        if lim != 3:
            assert O.L.wxo_average(lim, a, -a) == 0.0
        b = rng.normal()
        if lim in (1, 2, 3):
            assert O.L.wxo_average(lim, a, b) == pytest.approx(O.L.wxo_average(lim, b, a), rel=1e-14)
    # minmod and van Leer closed forms
    for a, b in rng.normal(0, 1, (300, 2)):
        mm = 0.0 if a * b <= 0 else math.copysign(min(abs(a), abs(b)), a)
        assert O.L.wxo_average(1, a, b) == pytest.approx(mm, rel=1e-15, abs=0)
        vl = 0.0 if a * b <= 0 else a * b * (a + b) / (a * a + b * b)
        assert O.L.wxo_average(2, a, b) == pytest.approx(vl, rel=1e-15, abs=0)
    assert O.L.wxo_average(-1, 1.0, 2.0) == 0.0
    assert O.L.wxo_average(0, 1.0, 2.0) == 1.5


def test_uprim_primu_roundtrip_and_floors():
    pp = cfg_ad().params()
    rng = np.random.default_rng(5)
    for q in prims(rng, 500):
        back, T = O.uprim(pp, O.primu(pp, q))
        np.testing.assert_allclose(back, q, rtol=1e-10, atol=1e-13)
        assert T == pytest.approx(q[3] / q[0], rel=1e-10)
    q, _ = O.uprim(pp, np.array([-1.0, 0.0, 0.0, -5.0, 0.0]))
    assert q[0] == 1e-15 and q[3] == 1e-15
    # COOLINGH: T >= 1 K and P recomputed from T (src/uprim.f90:37-41)
    ph = P.config_c3().params()
    n, y, T = 100.0, 0.3, 0.5
    uu = np.array([n, 0.0, 0.0, 1.5 * n * (2 - y) * T / ph.Tempsc, n * y])
    q, Th = O.uprim(ph, uu)
    assert Th == 1.0 and q[3] == pytest.approx(n * (2 - y) / ph.Tempsc, rel=1e-15)


def test_dmc_table_and_interpolation():
    pp = P.config_c2().params()
    tab = np.array(pp.cooltab[:]).reshape(41, 2)
    np.testing.assert_allclose(np.log10(tab[:, 0]), 4.0 + 0.1 * np.arange(41), atol=1e-12)
    assert tab[0, 1] == pytest.approx(10 ** -25.407, rel=1e-14)     # first row of coolingDMC.tab
    assert tab[40, 1] == pytest.approx(10 ** -22.680, rel=1e-14)    # last row
    for i in range(1, 40):
        Tn = tab[i, 0] * (1 + 1e-12)
        assert O.cooldmc(pp, Tn) == pytest.approx(tab[i, 1], rel=1e-9)
        Tm = 0.5 * (tab[i, 0] + tab[i + 1, 0])
        assert O.cooldmc(pp, Tm) == pytest.approx(0.5 * (tab[i, 1] + tab[i + 1, 1]), rel=1e-9)
    assert O.cooldmc(pp, 4e8) == pytest.approx(0.21e-26 * math.sqrt(4e8), rel=1e-15)


def test_neutral_fraction_solves_the_ionisation_ode():
    """coolingh's y1 (src/coolingh.f90:218-228) is the exact solution of
    dy/dt = n [a y^2 + b y + c] with a = alpha+C, b = -((2+xi)alpha+(1+xi)C),
    c = (1+xi)alpha (recombination minus collisional ionisation at n_e =
    n(1-y+xi)); checked against a high-accuracy ODE integration."""
    from scipy.integrate import solve_ivp
    pp = P.config_c3().params()
    rng = np.random.default_rng(6)
    xi = 1e-4
    checked = 0
    for _ in range(60):
        n = 10 ** rng.uniform(0, 3)
        T = 10 ** rng.uniform(3.8, 4.6)
        y0 = rng.uniform(0.05, 0.95)
        alpha = 2.55e-13 * (1e4 / T) ** 0.79
        C = 5.83e-11 * math.sqrt(T) * math.exp(-157828. / T)
        a, b, c = alpha + C, -((2 + xi) * alpha + (1 + xi) * C), (1 + xi) * alpha
        dt = 0.3 / (n * (abs(b) + 1e-30))
        sol = solve_ivp(lambda t, y: n * (a * y * y + b * y + c), (0, dt), [y0], rtol=1e-12, atol=1e-14,
                        method="DOP853")
        yode = sol.y[0, -1]
        if not (1e-4 < yode < 1.0):
            continue
        uu = np.array([n, 0.0, 0.0, 1.5 * n * (2 - y0) * T / pp.Tempsc, n * y0])
        out = O.atomic(pp, dt, uu)
        assert out[4] / n == pytest.approx(yode, rel=1e-7)
        checked += 1
    assert checked > 30


def test_uniform_state_is_preserved_exactly():
    cfg = cfg_ad(bc=("OUTFLOW",) * 4, nlevs=1, ic="ambient", ic_args=dict(n=2.0, T=1.0, yneutral=0.3))
    pp, m = oracle_mesh(O, cfg, 64)
    m.u[1:17, 1] = 0.7 * m.u[1:17, 0]
    m.u[1:17, 3] += 0.5 * 0.49 * m.u[1:17, 0]
    for nb in range(1, 17):
        m.primit[nb] = O.calcprim(pp, m.u[nb])
    u0 = m.u[1:17, :, 2:18, 2:18].copy()
    t = 0.0
    for _ in range(5):
        dt = O.timestep(pp, m)
        t += dt
        O.tstep(pp, m, t, dt)
    assert np.array_equal(m.u[1:17, :, 2:18, 2:18], u0)


def test_mass_energy_conserved_on_closed_single_level_box():
    cfg = cfg_ad(nlevs=1)
    pp, m = oracle_mesh(O, cfg, 64)
    lv = m.leaves()
    dx, dy = pp.dx[0], pp.dy[0]

    def totals():
        return (m.u[lv, 0, 2:18, 2:18].sum() * dx * dy, m.u[lv, 3, 2:18, 2:18].sum() * dx * dy)
    M0, E0 = totals()
    t = 0.0
    for _ in range(30):
        dt = O.timestep(pp, m)
        t += dt
        O.tstep(pp, m, t, dt)
    M1, E1 = totals()
    assert abs(M1 - M0) <= 1e-13 * M0
    assert abs(E1 - E0) <= 1e-12 * E0


def _random_amr(seed, cfg=None, nsteps=3):
    cfg = cfg or cfg_ad(nlevs=4, bc=("OUTFLOW", "REFLECTIVE", "OUTFLOW", "REFLECTIVE"))
    pp, m = oracle_mesh(O, cfg, 3000)
    rng = np.random.default_rng(seed)
    for nb in range(1, m.nbmax + 1):
        m.u[nb, 0] *= 1 + 0.5 * rng.uniform(0, 1, (20, 20)) * (rng.uniform() < 0.4)
        m.primit[nb] = O.calcprim(pp, m.u[nb])
    for _ in range(nsteps):
        O.admesh(pp, m)
    return cfg, pp, m


def _geometric_neighbour(m, nb, face):
    """Neighbour per the lp convention, by geometry: the block at the same
    level adjacent across `face`, else the coarser block covering that spot."""
    lev = m.lp[nb, 0]
    ix, iy = m.icoord[nb]
    step = 16
    tx, ty = ix + (-step, step, 0, 0)[face], iy + (0, 0, -step, step)[face]
    for L in range(lev, 0, -1):
        sh = lev - L
        cx, cy = (tx >> sh) & ~15, (ty >> sh) & ~15
        for b in range(1, m.nbmax + 1):
            if m.lp[b, 0] == L and m.icoord[b, 0] == cx and m.icoord[b, 1] == cy and b != nb:
                return b
    return -1


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tree_pointers_match_geometry(seed):
    cfg, pp, m = _random_amr(seed)
    nx = cfg.nbrootx * 16
    ny = cfg.nbrooty * 16
    lv = m.leaves()
    assert len(lv) > 16
    for nb in lv:
        lev = m.lp[nb, 0]
        for f in range(4):
            ix, iy = m.icoord[nb]
            scale = 2 ** (lev - 1)
            edge = (ix == 0, ix + 16 == nx * scale, iy == 0, iy + 16 == ny * scale)[f]
            want = -1 if edge else _geometric_neighbour(m, nb, f)
            assert m.lp[nb, 4 + f] == want, (nb, f)
        # 2:1 balance across faces
        for f in range(4):
            n1 = m.lp[nb, 4 + f]
            if n1 > 0:
                assert abs(m.lp[n1, 0] - lev) <= 1


@pytest.mark.parametrize("seed", [1, 2])
def test_locatebounds_covers_every_leaf_face_once(seed):
    cfg, pp, m = _random_amr(seed)
    n = O.locatebounds(pp, m)
    cover = {}
    opp = {0: 1, 1: 0, 2: 3, 3: 2}
    for nbs, n1, code in m.bounds[:n]:
        if code < 0:
            cover[(nbs, -code - 1)] = cover.get((nbs, -code - 1), 0) + 1
        else:
            f = (code - 1) // 3
            kind = (code - 1) % 3
            if kind == 0:
                cover[(nbs, f)] = cover.get((nbs, f), 0) + 1
                cover[(n1, opp[f])] = cover.get((n1, opp[f]), 0) + 1
            else:
                if kind == 1:
                    cover[(nbs, f)] = cover.get((nbs, f), 0) + 1
                cover[(n1, opp[f])] = cover.get((n1, opp[f]), 0) + 1
                assert m.lp[n1, 0] == m.lp[nbs, 0] + 1
    for nb in m.leaves():
        for f in range(4):
            assert cover.get((nb, f), 0) == 1, (nb, f)


def test_refine_injection_and_coarse_mean():
    cfg = cfg_ad()
    pp, m = oracle_mesh(O, cfg, 200)
    rng = np.random.default_rng(9)
    m.u[1] = rng.uniform(1, 2, m.u[1].shape)
    up = np.zeros(m.cap + 1, dtype=np.int32)
    dn = np.zeros(m.cap + 1, dtype=np.int32)
    s = m.nbmax + 1
    m.lp[1, 2] = s
    m.s.nbmax = m.nbmax + 4
    O.refine(pp, m, 1, s, s + 1, s + 2, s + 3)
    O.updatelpup(pp, m, 1, s, s + 1, s + 2, s + 3, up, dn)
    for k, (ox, oy) in enumerate([(0, 0), (8, 0), (0, 8), (8, 8)]):
        son = m.u[s + k]
        for j in range(1, 17):
            for i in range(1, 17):
                assert np.array_equal(son[:, j + 1, i + 1], m.u[1][:, (j + 1) // 2 + oy + 1, (i + 1) // 2 + ox + 1])
    # perturb sons and coarsen: father interior = mean of 2x2, mass conserved
    for k in range(4):
        m.u[s + k] = rng.uniform(1, 2, m.u[s + k].shape)
    O.coarse(pp, m, 1, s, s + 1, s + 2, s + 3)
    fine = sum(m.u[s + k, 0, 2:18, 2:18].sum() for k in range(4)) * 0.25
    assert m.u[1, 0, 2:18, 2:18].sum() == pytest.approx(fine, rel=1e-14)
    assert m.u[1, 0, 2, 2] == pytest.approx(0.25 * m.u[s, 0, 2:4, 2:4].sum(), rel=1e-15)


def test_basegrid_matches_root_layout():
    cfg = P.config_c3(nbrootx=5, nbrooty=3)
    pp = cfg.params()
    m = MeshState(100, cfg.neq)
    O.basegrid(pp, m)
    lp, ic = P.root_layout(cfg)
    assert np.array_equal(m.lp[1:16], lp[1:]) and np.array_equal(m.icoord[1:16], ic[1:])


def test_timestep_is_the_cfl_minimum():
    cfg = cfg_ad(nlevs=1)
    pp, m = oracle_mesh(O, cfg, 64)
    dtmin = 1e30
    for nb in m.leaves():
        q = m.primit[nb][:, 2:18, 2:18]
        cs = np.sqrt(pp.gamma * q[3] / q[0])
        dtmin = min(dtmin, (pp.dx[0] / (np.abs(q[1]) + cs)).min(), (pp.dy[0] / (np.abs(q[2]) + cs)).min())
    assert O.timestep(pp, m) == pytest.approx(pp.cfl * dtmin, rel=1e-15)
