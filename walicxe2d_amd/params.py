"""Run-time parameters and problem set-ups for the Walicxe-2D hot path.

The reference fixes every physics/BC/solver choice at compile time through
cpp switches (JET/Makefile:27-62 -> 118-209) and `parameters.f90` constants
(src/parameters.f90:12-89).  Here they are one run-time structure,
`WxParams`, whose binary layout is the C-ABI's `wx_params`
(include/wx_capi.h).

Scalings follow src/parameters.f90:48-54 with the same evaluation order, and
grid spacings follow src/initmain.f90:58-62 generalised to an
nbrootx x nbrooty root layout.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

NX = 16
NY = 16
NP = 20            # padded extent, nghost = 2
NCELL = NP * NP
MAXLEV = 16

SOLVERS = {"HLLC": 0, "HLL": 1, "HLL_HLLC": 2}
COOLING = {"ADIABATIC": 0, "H": 1, "DMC": 2}
BCS = {"OUTFLOW": 0, "REFLECTIVE": 1}

# bit 0: HLLC/hybrid right-star energy typo (src/hllcfluxes.f90:157,
#        src/hll_hllcfluxes.f90:192); bit 1: OUTFYB writes block `nb` = list
#        position (src/boundaryI.f90:87).
COMPAT_HLLC_TYPO = 1
COMPAT_OUTFYB = 2
COMPAT_ALL = COMPAT_HLLC_TYPO | COMPAT_OUTFYB

# physical constants, src/parameters.f90:28-34
PC = 3.0857E18
AMH = 1.66e-24
RG = 8.3145e7
YR = 3.156E7
MSUN = 1.99E33
T0 = 1.e4


class WxParams(ctypes.Structure):
    """Binary twin of `wx_params` (include/wx_capi.h)."""
    _fields_ = [
        ("neq", ctypes.c_int32), ("nlevs", ctypes.c_int32),
        ("nbrootx", ctypes.c_int32), ("nbrooty", ctypes.c_int32),
        ("solver", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("cooling", ctypes.c_int32), ("otherb", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 4), ("cylindrical", ctypes.c_int32),
        ("compat", ctypes.c_int32), ("nblocks", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("cv", ctypes.c_double), ("gamma", ctypes.c_double),
        ("eta", ctypes.c_double), ("cfl", ctypes.c_double),
        ("xmax", ctypes.c_double), ("ymax", ctypes.c_double),
        ("Tempsc", ctypes.c_double), ("Psc", ctypes.c_double), ("tsc", ctypes.c_double),
        ("tol", ctypes.c_double), ("tol2", ctypes.c_double),
        ("dx", ctypes.c_double * MAXLEV), ("dy", ctypes.c_double * MAXLEV),
        ("cooltab", ctypes.c_double * 82),
        ("jet", ctypes.c_double * 12),
    ]


def scalings(amol: float, xphys: float, xmax: float) -> dict:
    """src/parameters.f90:39-54, same operation order."""
    cv = 1.5
    gamma = (cv + 1.) / cv
    rsc = xphys / xmax
    rhosc = AMH * amol
    tempsc = T0 * gamma
    psc = gamma * AMH * RG * T0
    vsc2 = gamma * T0 * RG / amol
    tsc = rsc / math.sqrt(vsc2)
    return dict(cv=cv, gamma=gamma, rsc=rsc, rhosc=rhosc, Tempsc=tempsc,
                Psc=psc, vsc2=vsc2, tsc=tsc)


def dmc_table(path: str | None = None) -> np.ndarray:
    """cooltab(2,41) as src/initmain.f90:83-93 fills it: (10**a, 10**-b).

    Returned flattened in Fortran column-major order (T0, L0, T1, L1, ...).
    """
    if path is None:
        path = os.path.join(os.path.dirname(__file__), "data", "coolingDMC.tab")
    rows = []
    with open(path) as fh:
        for line in fh:
            parts = line.replace("d", "e").replace("D", "e").split()
            if len(parts) >= 2:
                rows.append((float(parts[0]), float(parts[1])))
    if len(rows) != 41:
        raise ValueError(f"DMC table {path}: expected 41 rows, got {len(rows)}")
    out = np.empty(82)
    for i, (a, b) in enumerate(rows):
        out[2 * i] = math.pow(10.0, a)
        out[2 * i + 1] = math.pow(10.0, -b)
    return out


@dataclass
class Config:
    """One problem set-up: the workloads the path is measured and checked on."""
    name: str
    neq: int
    nlevs: int
    nbrootx: int
    nbrooty: int
    xmax: float
    ymax: float
    xphys: float
    amol: float
    solver: str = "HLLC"
    limiter: int = 1
    cooling: str = "ADIABATIC"
    bc: tuple = ("OUTFLOW", "OUTFLOW", "OUTFLOW", "OUTFLOW")
    otherb: bool = False
    cylindrical: bool = False
    compat: int = COMPAT_ALL
    cfl: float = 0.4
    eta: float = 0.001
    tol: float = 0.05
    tol2: float = 1.0
    nblocks: int = 4096
    ic: str = "explosion"
    ic_args: dict = field(default_factory=dict)
    jet: dict = field(default_factory=dict)

    @property
    def sc(self) -> dict:
        return scalings(self.amol, self.xphys, self.xmax)

    def dxdy(self):
        dx = [self.xmax / NX / (2. ** (l - 1)) / float(self.nbrootx) for l in range(1, self.nlevs + 1)]
        dy = [self.ymax / NY / (2. ** (l - 1)) / float(self.nbrooty) for l in range(1, self.nlevs + 1)]
        return dx, dy

    def params(self) -> WxParams:
        sc = self.sc
        p = WxParams()
        p.neq = self.neq
        p.nlevs = self.nlevs
        p.nbrootx = self.nbrootx
        p.nbrooty = self.nbrooty
        p.solver = SOLVERS[self.solver]
        p.limiter = self.limiter
        p.cooling = COOLING[self.cooling]
        p.otherb = 1 if self.otherb else 0
        for k in range(4):
            p.bc[k] = BCS[self.bc[k]]
        p.cylindrical = 1 if self.cylindrical else 0
        p.compat = self.compat
        p.nblocks = self.nblocks
        p.cv = sc["cv"]
        p.gamma = sc["gamma"]
        p.eta = self.eta
        p.cfl = self.cfl
        p.xmax = self.xmax
        p.ymax = self.ymax
        p.Tempsc = sc["Tempsc"]
        p.Psc = sc["Psc"]
        p.tsc = sc["tsc"]
        p.tol = self.tol
        p.tol2 = self.tol2
        dx, dy = self.dxdy()
        for l in range(self.nlevs):
            p.dx[l] = dx[l]
            p.dy[l] = dy[l]
        if self.cooling == "DMC":
            tab = dmc_table()
            for k in range(82):
                p.cooltab[k] = tab[k]
        if self.otherb:
            j = self.jet_vector()
            for k in range(12):
                p.jet[k] = j[k]
        return p

    def jet_vector(self) -> list:
        """(RJ, LJ, posx, posy, denJ, TempJ, vJ0, dvJ, tau, y_neutral, 2-y, 0)
        in code units; the form of JET/jet.f90:15-35."""
        sc = self.sc
        j = self.jet
        y = j.get("yneutral", 1.e-4)
        return [j["RJ"], j["LJ"], j.get("posx", 0.), j.get("posy", 0.), j["den"],
                j["T"] / sc["Tempsc"], j["v0"] / math.sqrt(sc["vsc2"]),
                j["dv"] / math.sqrt(sc["vsc2"]), j["tau_yr"] * YR / sc["tsc"],
                y, 2. - y, 0.]


# --------------------------------------------------------------------------
# The workloads: C1 point explosion (adiabatic), C2 supernova with DMC
# cooling and 5 levels, C3 cooling jet, C4 explosion ensemble, C5 jet variants.
def config_c1(**kw) -> Config:
    """2-D point explosion, code units, 4x4 roots, AMR off, adiabatic."""
    c = Config(name="c1_explosion", neq=4, nlevs=1, nbrootx=4, nbrooty=4,
               xmax=1.0, ymax=1.0, xphys=PC, amol=0.6, cooling="ADIABATIC",
               bc=("OUTFLOW",) * 4, ic="explosion_code",
               ic_args=dict(rho=1.0, p=1e-5, E=1.0, r=0.02, xc=0.5, yc=0.5))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config_c2(**kw) -> Config:
    """SN-style explosion, 40 pc box, 8x8 roots, 5 levels, DMC cooling."""
    c = Config(name="c2_sn_dmc", neq=4, nlevs=5, nbrootx=8, nbrooty=8,
               xmax=1.0, ymax=1.0, xphys=40. * PC, amol=0.6, cooling="DMC",
               bc=("OUTFLOW",) * 4, tol=0.1, nblocks=40000, ic="sn",
               ic_args=dict(n=1.0, T=1.e4, E=1.e51, R=1.0 * PC, xc=0.5, yc=0.5))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config_c3(**kw) -> Config:
    """Variable slab jet, box 4x1 (x 1e17 cm), 64x16 roots, 4 levels, coolingh."""
    c = Config(name="c3_jet_coolingh", neq=5, nlevs=4, nbrootx=64, nbrooty=16,
               xmax=4.0, ymax=1.0, xphys=4.e17, amol=1.0, cooling="H",
               bc=("REFLECTIVE", "OUTFLOW", "REFLECTIVE", "OUTFLOW"), otherb=True,
               nblocks=60000, ic="ambient",
               ic_args=dict(n=100., T=1.e3, yneutral=0.9999),
               jet=dict(RJ=0.05, LJ=0.05, posx=0., posy=0., den=100., T=1.e3,
                        v0=200.e5, dv=50.e5, tau_yr=30., yneutral=1.e-4))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config_c4(i: int, **kw) -> Config:
    """Instance i (0..63) of the batched explosion ensemble: a C2-type
    supernova in a 40 pc box with 4x4 roots and DMC cooling, with per-instance
    draws from seed 20261202 + i in this order: log10 E ~ N(51, 0.5),
    log10 n ~ U(-1, 1), centre x, y ~ U(0.25, 0.75) (box units),
    max level ~ U{3, 4, 5, 6}."""
    rng = np.random.default_rng(20261202 + i)
    loge = rng.normal(51.0, 0.5)
    logn = rng.uniform(-1.0, 1.0)
    xc = rng.uniform(0.25, 0.75)
    yc = rng.uniform(0.25, 0.75)
    lmax = int(rng.integers(3, 7))
    c = Config(name=f"c4_sn_{i}", neq=4, nlevs=lmax, nbrootx=4, nbrooty=4,
               xmax=1.0, ymax=1.0, xphys=40. * PC, amol=0.6, cooling="DMC",
               bc=("OUTFLOW",) * 4, tol=0.1, nblocks={3: 1200, 4: 3000, 5: 8000, 6: 16000}[lmax], ic="sn",
               ic_args=dict(n=10.0 ** logn, T=1.e4, E=10.0 ** loge, R=1.0 * PC, xc=xc, yc=yc))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config_c5(k: int, **kw) -> Config:
    """Jet variant k of the cooling-surrogate data set: the C3 jet with seed
    777 + k drawing v0 ~ U(150, 250) km/s, then log10 n ~ U(log10 30,
    log10 300) cm^-3 for both the ambient medium and the jet; the velocity
    amplitude stays a quarter of v0."""
    rng = np.random.default_rng(777 + k)
    v0 = rng.uniform(150.e5, 250.e5)
    n = 10.0 ** rng.uniform(math.log10(30.), math.log10(300.))
    c = config_c3()
    c.name = f"c5_jet_{k}"
    c.ic_args = dict(c.ic_args, n=n)
    c.jet = dict(c.jet, v0=v0, dv=0.25 * v0, den=n)
    for key, v in kw.items():
        setattr(c, key, v)
    return c


CONFIGS = {"c1": config_c1, "c2": config_c2, "c3": config_c3}


# --------------------------------------------------------------------------
# Initial conditions over a block's padded range (the reference applies
# initial_conditions over nxmin:nxmax, src/user_mod.f90:39-66).
def cell_centres(cfg: Config, level: int, icoord: tuple):
    dx, dy = cfg.dxdy()
    ii = np.arange(-1, NX + 3, dtype=np.float64)
    x = ((ii + icoord[0]) - 0.5) * dx[level - 1]
    y = ((ii + icoord[1]) - 0.5) * dy[level - 1]
    return np.meshgrid(x, y, indexing="xy")   # arrays [j][i]


def initial_block(cfg: Config, level: int, icoord: tuple) -> np.ndarray:
    """Conserved variables [neq][20][20] (j, i) of one block at t=0."""
    sc = cfg.sc
    X, Y = cell_centres(cfg, level, icoord)
    u = np.zeros((cfg.neq, NP, NP))
    a = cfg.ic_args
    cv = sc["cv"]
    if cfg.ic == "explosion_code":
        # uniform rho, v=0, P; energy E spread uniformly over the cells whose
        # centres lie within r of (xc, yc) (cell counting on the finest grid)
        dx, dy = cfg.dxdy()
        r2 = (X - a["xc"]) ** 2 + (Y - a["yc"]) ** 2
        inside = r2 < a["r"] ** 2
        ncell = _explosion_cells(cfg)
        u[0] = a["rho"]
        u[3] = cv * a["p"]
        u[3][inside] += a["E"] / (ncell * dx[0] * dy[0])
    elif cfg.ic == "sn":
        n = a["n"]
        P = n * a["T"] / sc["Tempsc"]
        u[0] = n
        u[3] = cv * P
        xr = (X - a["xc"]) * sc["rsc"]
        yr = (Y - a["yc"]) * sc["rsc"]
        r = np.sqrt(xr ** 2 + yr ** 2)
        inside = r <= a["R"]
        eth = (3. / 4. / math.pi) * a["E"] / (a["R"] ** 3) / sc["Psc"]
        u[3][inside] = eth
    elif cfg.ic == "ambient":
        n = a["n"]
        y = a["yneutral"]
        u[0] = n
        u[3] = cv * (2. - y) * n * a["T"] / sc["Tempsc"]
        if cfg.neq >= 5:
            u[4] = n * y
        if cfg.neq >= 6:
            u[5] = -n
    else:
        raise ValueError(cfg.ic)
    return u


def _explosion_cells(cfg: Config) -> int:
    dx, dy = cfg.dxdy()
    nxg = cfg.nbrootx * NX
    nyg = cfg.nbrooty * NY
    xc = (np.arange(nxg) + 0.5) * dx[0]
    yc = (np.arange(nyg) + 0.5) * dy[0]
    X, Y = np.meshgrid(xc, yc)
    a = cfg.ic_args
    return int(np.count_nonzero((X - a["xc"]) ** 2 + (Y - a["yc"]) ** 2 < a["r"] ** 2))


def root_layout(cfg: Config):
    """Generalised root blocks at level 1, x fastest; the 1-D form is
    src/basegrid.f90:26-97 (lp(nb,1:8) = level, father, first son, sibling
    ID, neighbours left/right/bottom/top with -1 at the box edge)."""
    nrx, nry = cfg.nbrootx, cfg.nbrooty
    n = nrx * nry
    lp = np.zeros((n + 1, 8), dtype=np.int32)
    icoord = np.zeros((n + 1, 2), dtype=np.int32)
    for iy in range(nry):
        for ix in range(nrx):
            nb = 1 + ix + iy * nrx
            lp[nb] = (1, 0, 0, 1,
                      nb - 1 if ix > 0 else -1,
                      nb + 1 if ix < nrx - 1 else -1,
                      nb - nrx if iy > 0 else -1,
                      nb + nrx if iy < nry - 1 else -1)
            icoord[nb] = (ix * NX, iy * NY)
    return lp, icoord


def initial_state(cfg: Config):
    """(u, lp, icoord) of the root mesh at t=0; u is [nroot+1][neq][20][20]."""
    lp, icoord = root_layout(cfg)
    n = lp.shape[0] - 1
    u = np.zeros((n + 1, cfg.neq, NP, NP))
    for nb in range(1, n + 1):
        u[nb] = initial_block(cfg, 1, tuple(icoord[nb]))
    return u, lp, icoord
