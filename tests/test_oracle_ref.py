"""The C oracle against the reference's own Fortran (oracle/_ref), bit for bit.

oracle/_ref is built from the reference's unmodified hot-path sources by
oracle/ref/Makefile where the reference tree is present; tests skip where it
is absent (tests/test_oracle_golden.py then checks the committed vectors).
"""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from wxoracle import MeshState, Oracle, RefLib, ref_available

from harness import oracle_mesh

O = Oracle()
_libs = {}


def ref(variant):
    if not ref_available(variant):
        pytest.skip(f"oracle/_ref/libwxref_{variant}.so not built (reference tree absent)")
    if variant not in _libs:
        _libs[variant] = RefLib(variant)
    return _libs[variant]


def cfg_for(variant):
    if variant == "c1":
        return P.config_c1()
    if variant == "c2":
        return P.config_c2(nblocks=4096, tol=0.05)   # criteria's own threshold, src/criteria.f90:11
    if variant == "c3":
        return P.config_c3(nblocks=4096, nbrootx=16, nbrooty=4)
    base = dict(name=variant, neq=5, nlevs=3, nbrootx=4, nbrooty=4, xmax=1.0, ymax=1.0, xphys=P.PC,
                amol=0.6, cooling="ADIABATIC", bc=("REFLECTIVE", "OUTFLOW", "OUTFLOW", "REFLECTIVE"),
                ic="explosion_code", ic_args=dict(rho=1.0, p=1e-2, E=1.0, r=0.1, xc=0.4, yc=0.45))
    if variant == "hll":
        base["solver"] = "HLL"
    elif variant == "hybrid":
        base["solver"] = "HLL_HLLC"
    elif variant.startswith("lim"):
        base["limiter"] = int(variant[3:])
    elif variant == "cyl":
        base["limiter"] = 2
        base["cylindrical"] = True
    return P.Config(**base)


def random_prims(rng, n, neq, vmax=3.0):
    q = np.empty((n, neq))
    q[:, 0] = 10 ** rng.uniform(-2, 2, n)
    q[:, 1] = rng.uniform(-vmax, vmax, n)
    q[:, 2] = rng.uniform(-vmax, vmax, n)
    q[:, 3] = 10 ** rng.uniform(-3, 2, n)
    for k in range(4, neq):
        q[:, k] = q[:, 0] * rng.uniform(0, 1, n)
    return q


@pytest.mark.parametrize("variant", ["c1", "c2", "c3"])
def test_constants(variant):
    R = ref(variant)
    sc = cfg_for(variant).sc
    for k in ("Tempsc", "Psc", "tsc", "gamma", "cv"):
        assert R.consts[k] == sc[k], k


@pytest.mark.parametrize("variant", ["c1", "c2", "c3"])
def test_uprim_primu_primf_bitexact(variant):
    R = ref(variant)
    cfg = cfg_for(variant)
    pp = cfg.params()
    rng = np.random.default_rng(1)
    for q in random_prims(rng, 500, cfg.neq):
        uu = O.primu(pp, q)
        assert np.array_equal(uu, R.primu(q))
        assert np.array_equal(O.primf(pp, q), R.primf(q))
        a, Ta = O.uprim(pp, uu)
        b, Tb = R.uprim(uu)
        assert np.array_equal(a, b) and Ta == Tb
    # floors: negative density / pressure
    for uu in ([-1.0, 0.0, 0.0, 1.0] + [0.1] * (cfg.neq - 4), [1.0, 2.0, 0.0, 0.1] + [0.5] * (cfg.neq - 4)):
        uu = np.array(uu)
        a, Ta = O.uprim(pp, uu)
        b, Tb = R.uprim(uu)
        assert np.array_equal(a, b) and Ta == Tb


@pytest.mark.parametrize("variant", ["lim0", "c1", "lim2", "lim3", "lim4", "lim5", "lim6"])
def test_limiter_bitexact(variant):
    R = ref(variant)
    pp = cfg_for(variant).params()
    rng = np.random.default_rng(2)
    neq = R.neq
    for _ in range(400):
        a = rng.normal(0, 1, (4, neq))
        a[:, rng.integers(0, neq)] = 0.0   # exact zeros exercise the sign branches
        pl_o, pr_o = O.limiter(pp, a[0], a[1], a[2], a[3])
        pl_r, pr_r = R.limiter(a[0], a[1], a[2], a[3])
        assert np.array_equal(pl_o, pl_r) and np.array_equal(pr_o, pr_r)


def _amr_mesh(variant, nref=2, seed=3):
    """A small multi-level mesh with random data (refines a few leaves)."""
    cfg = cfg_for(variant)
    pp, m = oracle_mesh(O, cfg, 600)
    rng = np.random.default_rng(seed)
    if cfg.nlevs > 1:
        for _ in range(nref):
            up = np.zeros(m.cap + 1, dtype=np.int32)
            dn = np.zeros(m.cap + 1, dtype=np.int32)
            leaves = m.leaves()
            pick = rng.choice(leaves, size=max(1, len(leaves) // 5), replace=False)
            for nb in pick:
                if m.lp[nb, 0] < cfg.nlevs:
                    up[nb] = 1
            # the oracle's admesh takes flags from criteria; emulate a forced mark
            for nb in sorted(pick):
                if m.lp[nb, 0] < cfg.nlevs and m.leaf[nb] and m.lp[nb, 2] == 0:
                    s = m.nbmax + 1
                    m.lp[nb, 2] = s
                    m.s.nbmax = m.nbmax + 4
                    O.refine(pp, m, nb, s, s + 1, s + 2, s + 3)
                    O.updatelpup(pp, m, nb, s, s + 1, s + 2, s + 3, up, dn)
    for nb in range(1, m.nbmax + 1):
        m.u[nb] *= 1 + 0.2 * rng.uniform(-1, 1, m.u[nb].shape)
        m.primit[nb] = O.calcprim(pp, m.u[nb])
    m.up[...] = m.u * (1 + 0.1 * rng.uniform(-1, 1, m.u.shape))
    return cfg, pp, m


@pytest.mark.parametrize("variant", ["c1", "hll", "hybrid", "lim2", "cyl"])
@pytest.mark.parametrize("choice", [1, 2])
def test_fluxes_bitexact(variant, choice):
    R = ref(variant)
    cfg, pp, m = _amr_mesh(variant)
    R.setup(pp)
    R.put(m)
    for nb in m.leaves()[:12]:
        f, g, err = O.fluxes(pp, m, nb, choice)
        fr, gr = R.fluxes(nb, choice)
        assert err == 0
        sl = np.s_[:, 1:18, 1:18]
        assert np.array_equal(f[sl], fr[sl]) and np.array_equal(g[sl], gr[sl]), (nb, choice)


def test_hllc_right_star_typo_branch():
    """States with S* < 0 and S* != u_R: the only branch the reference's typo
    (src/hllcfluxes.f90:157) reaches; pinned by the reference itself.  The
    same fluxes with the typo switched off must differ, proving the branch
    was exercised."""
    R = ref("c1")
    cfg = cfg_for("c1")
    pp = cfg.params()
    pfix = cfg_for("c1").params()
    pfix.compat = 0
    m = MeshState(32, 4)
    O.basegrid(pp, m)
    rng = np.random.default_rng(5)
    differs = False
    for trial in range(4):
        for nb in range(1, m.nbmax + 1):
            rho = 10 ** rng.uniform(-1, 1, (20, 20))
            vx = rng.uniform(-2, 0.5, (20, 20))
            pr = 10 ** rng.uniform(-2, 0, (20, 20))
            m.u[nb] = np.stack([rho, rho * vx, rho * rng.uniform(-1, 1, (20, 20)), 1.5 * pr + 0.5 * rho * vx ** 2])
            m.primit[nb] = O.calcprim(pp, m.u[nb])
        R.put(m)
        for nb in range(1, m.nbmax + 1):
            f, g, err = O.fluxes(pp, m, nb, 1)
            assert err == 0
            fr, gr = R.fluxes(nb, 1)
            assert np.array_equal(f, fr) and np.array_equal(g, gr)
            f0, g0, _ = O.fluxes(pfix, m, nb, 1)
            differs |= not np.array_equal(f0, f)
    assert differs


@pytest.mark.parametrize("variant", ["c1", "c3", "hll"])
@pytest.mark.parametrize("which", [1, 2])
def test_boundaries_bitexact(variant, which):
    """boundaryI / boundaryII over an AMR mesh, whole padded blocks (corners too:
    the oracle applies the list in the reference's order)."""
    R = ref(variant)
    cfg, pp, m = _amr_mesh(variant)
    R.setup(pp)
    O.locatebounds(pp, m)
    R.put(m)
    t = 0.0123
    O.boundary(pp, m, t, which)
    R.L.wxr_boundary(t, which)
    m2 = m.copy()
    R.get(m2)
    lv = list(range(1, m.nbmax + 1))
    arr, arr2 = (m.up, m2.up) if which == 2 else (m.u, m2.u)
    assert np.array_equal(arr[lv], arr2[lv])


def test_atomic_aloss_bitexact():
    R = ref("c3")
    cfg = cfg_for("c3")
    pp = cfg.params()
    rng = np.random.default_rng(7)
    for _ in range(600):
        n = 10 ** rng.uniform(-1, 4)
        T = 10 ** rng.uniform(1, 8)
        y = rng.uniform(0, 1)
        dt = 10 ** rng.uniform(2, 11)
        vx, vy = rng.normal(0, 1, 2)
        Pth = n * (2 - y) * T / pp.Tempsc
        uu = np.array([n, n * vx, n * vy, 1.5 * Pth + 0.5 * n * (vx * vx + vy * vy), n * y])
        assert np.array_equal(O.atomic(pp, dt, uu), R.atomic(dt, uu))
        args = (rng.uniform(0, 1), rng.uniform(0, 1), dt, n, n * y, T)
        assert O.aloss(*args) == R.aloss(*args)


def test_cooldmc_bitexact():
    R = ref("c2")
    pp = cfg_for("c2").params()
    R.setup(pp)
    for T in np.concatenate([10 ** np.linspace(4.0001, 7.9999, 400), [2e8, 1e9]]):
        assert O.cooldmc(pp, T) == R.cooldmc(T)


@pytest.mark.parametrize("variant,steps", [("c1", 4), ("c2", 4), ("c3", 6), ("cyl", 3), ("hybrid", 3)])
def test_tstep_timestep_bitexact(variant, steps):
    """Whole tstep + timestep (incl. admesh between steps on the oracle side,
    replayed into the reference) on identical inputs: every cell equal."""
    R = ref(variant)
    cfg = cfg_for(variant)
    pp, m = oracle_mesh(O, cfg, 3000)
    R.setup(pp)
    if cfg.nlevs > 1:
        for _ in range(3):
            O.admesh(pp, m)
    t = 0.0
    for s in range(steps):
        O.locatebounds(pp, m)
        R.put(m)
        dt_o = O.timestep(pp, m)
        dt_r = R.L.wxr_timestep()
        assert dt_o == dt_r
        t += dt_o
        assert O.tstep(pp, m, t, dt_o) == 0
        R.L.wxr_tstep(t, dt_o)
        m2 = m.copy()
        R.get(m2)
        lv = m.leaves()
        assert np.array_equal(m.u[lv], m2.u[lv]), (variant, s)
        assert np.array_equal(m.primit[lv], m2.primit[lv]), (variant, s)
        if cfg.nlevs > 1:
            assert O.admesh(pp, m) == 0


@pytest.mark.parametrize("variant", ["c2", "c3"])
def test_criteria_refine_coarse_bitexact(variant):
    R = ref(variant)
    cfg = cfg_for(variant)
    pp, m = oracle_mesh(O, cfg, 3000)
    R.setup(pp)
    for _ in range(2):
        O.admesh(pp, m)
    R.put(m)
    up_o = np.zeros(m.cap + 1, dtype=np.int32)
    dn_o = np.zeros(m.cap + 1, dtype=np.int32)
    up_r, dn_r = R.flags()
    for nb in m.leaves():
        O.criteria(pp, m, nb, up_o, dn_o)
        R.L.wxr_criteria(nb, up_r.ctypes.data_as(type(m.s.lp)), dn_r.ctypes.data_as(type(m.s.lp)))
    assert np.array_equal(up_o[:m.cap + 1], up_r[:m.cap + 1])
    assert np.array_equal(dn_o[:m.cap + 1], dn_r[:m.cap + 1])
    # refine one leaf with both implementations (data + pointer surgery)
    nb = next(b for b in m.leaves() if m.lp[b, 0] < cfg.nlevs)
    s = m.nbmax + 1
    m.lp[nb, 2] = s
    m.s.nbmax = m.nbmax + 4
    R.put(m)
    O.refine(pp, m, nb, s, s + 1, s + 2, s + 3)
    O.updatelpup(pp, m, nb, s, s + 1, s + 2, s + 3, up_o, dn_o)
    R.L.wxr_refine(nb, s, s + 1, s + 2, s + 3, up_r.ctypes.data_as(type(m.s.lp)), dn_r.ctypes.data_as(type(m.s.lp)))
    m2 = m.copy()
    R.get(m2)
    ids = list(range(1, m.nbmax + 1))
    assert np.array_equal(m.lp[ids], m2.lp[ids])
    assert np.array_equal(m.icoord[ids], m2.icoord[ids])
    assert np.array_equal(m.u[ids], m2.u[ids]) and np.array_equal(m.primit[ids], m2.primit[ids])
    # coarsen the four sons back into their father
    R.put(m)
    O.coarse(pp, m, nb, s, s + 1, s + 2, s + 3)
    R.L.wxr_coarse(nb, s, s + 1, s + 2, s + 3)
    m3 = m.copy()
    R.get(m3)
    assert np.array_equal(m.u[nb], m3.u[nb]) and np.array_equal(m.primit[nb], m3.primit[nb])
