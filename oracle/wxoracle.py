"""ctypes front-end for the CPU oracles.  TEST INFRASTRUCTURE ONLY.

* `Oracle`  -- the C restatement (oracle/libwxoracle.so, built from
  oracle/wxoracle.c);
* `RefLib`  -- the reference's own hot-path Fortran, compiled unmodified into
  oracle/_ref/libwxref_<variant>.so by oracle/ref/Makefile (present only where
  the reference tree was available at build time).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NP = 20
NCELL = 400

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


class WxoMesh(ctypes.Structure):
    _fields_ = [("cap", ctypes.c_int32), ("nbmax", ctypes.c_int32),
                ("nbleafs", ctypes.c_int32), ("ninternal", ctypes.c_int32),
                ("u", _dp), ("up", _dp), ("primit", _dp),
                ("lp", _ip), ("leaf", _ip), ("icoord", _ip), ("bounds", _ip)]


class MeshState:
    """Host arrays of an AMR mesh in the oracle layout [id][var][j][i]."""

    def __init__(self, cap: int, neq: int):
        self.cap = cap
        self.neq = neq
        n = cap + 1
        self.u = np.zeros((n, neq, NP, NP))
        self.up = np.zeros((n, neq, NP, NP))
        self.primit = np.zeros((n, neq, NP, NP))
        self.lp = np.zeros((n, 8), dtype=np.int32)
        self.leaf = np.zeros(n, dtype=np.int32)
        self.icoord = np.zeros((n, 2), dtype=np.int32)
        self.bounds = np.zeros((4 * n, 3), dtype=np.int32)
        self.s = WxoMesh(cap, 0, 0, 0, dptr(self.u), dptr(self.up), dptr(self.primit),
                         iptr(self.lp), iptr(self.leaf), iptr(self.icoord), iptr(self.bounds))

    @property
    def nbmax(self):
        return self.s.nbmax

    @property
    def nbleafs(self):
        return self.s.nbleafs

    @property
    def ninternal(self):
        return self.s.ninternal

    def leaves(self):
        return [nb for nb in range(1, self.s.nbmax + 1) if self.leaf[nb]]

    def copy(self) -> "MeshState":
        m = MeshState(self.cap, self.neq)
        for k in ("u", "up", "primit", "lp", "leaf", "icoord", "bounds"):
            getattr(m, k)[...] = getattr(self, k)
        m.s.nbmax, m.s.nbleafs, m.s.ninternal = self.s.nbmax, self.s.nbleafs, self.s.ninternal
        return m


class Oracle:
    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "libwxoracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path}: build it with `make -C oracle`")
        L = ctypes.CDLL(path)
        self.L = L
        P = ctypes.c_void_p
        M = ctypes.POINTER(WxoMesh)
        L.wxo_uprim.argtypes = [P, _dp, _dp, _dp]
        L.wxo_primu.argtypes = [P, _dp, _dp]
        L.wxo_primf.argtypes = [P, _dp, _dp]
        L.wxo_average.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.wxo_average.restype = ctypes.c_double
        L.wxo_limiter.argtypes = [P, _dp, _dp, _dp, _dp]
        L.wxo_riemann.argtypes = [P, _dp, _dp, _dp]
        L.wxo_fluxes.argtypes = [P, M, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.wxo_step.argtypes = [P, M, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.wxo_viscosity.argtypes = [P, M, ctypes.c_int]
        L.wxo_calcprim.argtypes = [P, _dp, _dp]
        L.wxo_boundary.argtypes = [P, M, ctypes.c_double, ctypes.c_int]
        L.wxo_aloss.argtypes = [ctypes.c_double] * 6
        L.wxo_aloss.restype = ctypes.c_double
        L.wxo_atomic.argtypes = [P, ctypes.c_double, _dp]
        L.wxo_cooldmc.argtypes = [P, ctypes.c_double]
        L.wxo_cooldmc.restype = ctypes.c_double
        L.wxo_coolingh.argtypes = [P, M, ctypes.c_double]
        L.wxo_coolingdmc.argtypes = [P, M, ctypes.c_double]
        L.wxo_timestep.argtypes = [P, M]
        L.wxo_timestep.restype = ctypes.c_double
        L.wxo_tstep.argtypes = [P, M, ctypes.c_double, ctypes.c_double]
        L.wxo_locatebounds.argtypes = [P, M]
        L.wxo_criteria.argtypes = [P, M, ctypes.c_int, _ip, _ip]
        L.wxo_refine.argtypes = [P, M] + [ctypes.c_int] * 5
        L.wxo_updatelpup.argtypes = [P, M] + [ctypes.c_int] * 5 + [_ip, _ip]
        L.wxo_coarse.argtypes = [P, M] + [ctypes.c_int] * 5
        L.wxo_updatelpdown.argtypes = [P, M] + [ctypes.c_int] * 5
        L.wxo_admesh.argtypes = [P, M]
        L.wxo_markref.argtypes = [P, M, _ip, _ip]
        L.wxo_admesh_flags.argtypes = [P, M, _ip, _ip]
        L.wxo_critneighup.argtypes = [P, M, _ip]
        L.wxo_critneighdown.argtypes = [P, M, _ip]
        L.wxo_basegrid.argtypes = [P, M]
        L.wxo_run.argtypes = [P, M, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.wxo_hilbert2d.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int]
        L.wxo_hilbert2d.restype = ctypes.c_int32
        L.wxo_indexx.argtypes = [ctypes.c_int, _ip, _ip]
        L.wxo_lb_plan.argtypes = [_ip, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip]

    @staticmethod
    def _p(params):
        return ctypes.cast(ctypes.byref(params), ctypes.c_void_p)

    def uprim(self, params, uu):
        uu = np.ascontiguousarray(uu, dtype=np.float64)
        prim = np.zeros_like(uu)
        T = np.zeros(1)
        self.L.wxo_uprim(self._p(params), dptr(prim), dptr(uu), dptr(T))
        return prim, T[0]

    def primu(self, params, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        out = np.zeros_like(prim)
        self.L.wxo_primu(self._p(params), dptr(prim), dptr(out))
        return out

    def primf(self, params, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        out = np.zeros_like(prim)
        self.L.wxo_primf(self._p(params), dptr(prim), dptr(out))
        return out

    def limiter(self, params, pll, pl, pr, prr):
        pl = np.array(pl, dtype=np.float64)
        pr = np.array(pr, dtype=np.float64)
        self.L.wxo_limiter(self._p(params), dptr(np.ascontiguousarray(pll, dtype=np.float64)),
                           dptr(pl), dptr(pr), dptr(np.ascontiguousarray(prr, dtype=np.float64)))
        return pl, pr

    def riemann(self, params, pl, pr):
        pl = np.ascontiguousarray(pl, dtype=np.float64)
        pr = np.ascontiguousarray(pr, dtype=np.float64)
        ff = np.zeros_like(pl)
        err = self.L.wxo_riemann(self._p(params), dptr(pl), dptr(pr), dptr(ff))
        return ff, err

    def fluxes(self, params, m: MeshState, nb, choice):
        f = np.zeros((m.neq, NP, NP))
        g = np.zeros((m.neq, NP, NP))
        err = self.L.wxo_fluxes(self._p(params), ctypes.byref(m.s), nb, choice, dptr(f), dptr(g))
        return f, g, err

    def step(self, params, m, nb, dt, f, g):
        self.L.wxo_step(self._p(params), ctypes.byref(m.s), nb, dt, dptr(f), dptr(g))

    def viscosity(self, params, m, nb):
        self.L.wxo_viscosity(self._p(params), ctypes.byref(m.s), nb)

    def calcprim(self, params, ublk):
        ublk = np.ascontiguousarray(ublk, dtype=np.float64)
        out = np.zeros_like(ublk)
        self.L.wxo_calcprim(self._p(params), dptr(ublk), dptr(out))
        return out

    def boundary(self, params, m, time, which):
        self.L.wxo_boundary(self._p(params), ctypes.byref(m.s), time, which)

    def aloss(self, *a):
        return self.L.wxo_aloss(*[float(x) for x in a])

    def atomic(self, params, dt, uu):
        uu = np.array(uu, dtype=np.float64)
        self.L.wxo_atomic(self._p(params), dt, dptr(uu))
        return uu

    def cooldmc(self, params, T):
        return self.L.wxo_cooldmc(self._p(params), float(T))

    def coolingh(self, params, m, dt):
        self.L.wxo_coolingh(self._p(params), ctypes.byref(m.s), dt)

    def coolingdmc(self, params, m, dt):
        self.L.wxo_coolingdmc(self._p(params), ctypes.byref(m.s), dt)

    def timestep(self, params, m):
        return self.L.wxo_timestep(self._p(params), ctypes.byref(m.s))

    def tstep(self, params, m, time, dt):
        return self.L.wxo_tstep(self._p(params), ctypes.byref(m.s), time, dt)

    def locatebounds(self, params, m):
        return self.L.wxo_locatebounds(self._p(params), ctypes.byref(m.s))

    def criteria(self, params, m, nb, up, dn):
        self.L.wxo_criteria(self._p(params), ctypes.byref(m.s), nb, iptr(up), iptr(dn))

    def refine(self, params, m, nb, s1, s2, s3, s4):
        self.L.wxo_refine(self._p(params), ctypes.byref(m.s), nb, s1, s2, s3, s4)

    def updatelpup(self, params, m, nb, s1, s2, s3, s4, up, dn):
        self.L.wxo_updatelpup(self._p(params), ctypes.byref(m.s), nb, s1, s2, s3, s4, iptr(up), iptr(dn))

    def coarse(self, params, m, dad, s1, s2, s3, s4):
        self.L.wxo_coarse(self._p(params), ctypes.byref(m.s), dad, s1, s2, s3, s4)

    def updatelpdown(self, params, m, dad, s1, s2, s3, s4):
        self.L.wxo_updatelpdown(self._p(params), ctypes.byref(m.s), dad, s1, s2, s3, s4)

    def admesh(self, params, m):
        return self.L.wxo_admesh(self._p(params), ctypes.byref(m.s))

    def markref(self, params, m):
        """src/markref.f90: refinement / coarsening flags (by block ID, 0..cap)."""
        up = np.zeros(m.cap + 1, dtype=np.int32)
        dn = np.zeros(m.cap + 1, dtype=np.int32)
        self.L.wxo_markref(self._p(params), ctypes.byref(m.s), iptr(up), iptr(dn))
        return up, dn

    def admesh_flags(self, params, m, up, dn):
        """admesh (src/admesh.f90:24-105) from given markref flags."""
        return self.L.wxo_admesh_flags(self._p(params), ctypes.byref(m.s), iptr(up), iptr(dn))

    def critneighup(self, params, m, up):
        self.L.wxo_critneighup(self._p(params), ctypes.byref(m.s), iptr(up))

    def critneighdown(self, params, m, dn):
        self.L.wxo_critneighdown(self._p(params), ctypes.byref(m.s), iptr(dn))

    def hilbert2d(self, x, y, bits):
        """src/loadbalance.f90:251-306: the Hilbert key of (x, y) over `bits` bits."""
        return int(self.L.wxo_hilbert2d(int(x), int(y), int(bits)))

    def indexx(self, arr):
        """src/loadbalance.f90:320-400: 1-based index table of arr in ascending order."""
        a = np.ascontiguousarray(arr, dtype=np.int32)
        ix = np.zeros(len(a), dtype=np.int32)
        self.L.wxo_indexx(len(a), iptr(a), iptr(ix))
        return ix

    def lb_plan(self, lp, leaf, icoord, nbmax, nlevs, nblocks, nranks):
        """src/loadbalance.f90:24-33,73-105 for blocks 1..nbmax held on one
        rank: (err, key[nbmax+1], idnew[nbmax+1]), slot 0 unused.  lp
        [>nbmax][8], leaf [>nbmax], icoord [>nbmax][2], indexed by block ID."""
        lp = np.ascontiguousarray(lp[:nbmax + 1], dtype=np.int32)
        leaf = np.ascontiguousarray(leaf[:nbmax + 1], dtype=np.int32)
        ic = np.ascontiguousarray(icoord[:nbmax + 1], dtype=np.int32)
        key = np.zeros(nbmax + 1, dtype=np.int32)
        idnew = np.zeros(nbmax + 1, dtype=np.int32)
        err = self.L.wxo_lb_plan(iptr(lp), iptr(leaf), iptr(ic), nbmax, nlevs, nblocks, nranks,
                                 iptr(key), iptr(idnew))
        return err, key, idnew

    def basegrid(self, params, m):
        self.L.wxo_basegrid(self._p(params), ctypes.byref(m.s))

    def run(self, params, m, nsteps, time, do_admesh=True):
        t = np.array([time])
        dtl = np.zeros(1)
        z = np.zeros(1)
        err = self.L.wxo_run(self._p(params), ctypes.byref(m.s), nsteps, dptr(t), dptr(dtl),
                             1 if do_admesh else 0, dptr(z))
        return t[0], dtl[0], z[0], err


def ref_path(variant: str) -> str:
    return os.path.join(HERE, "_ref", f"libwxref_{variant}.so")


def ref_available(variant: str) -> bool:
    return os.path.exists(ref_path(variant))


class RefLib:
    """The reference's own Fortran routines behind the oracle/ref glue.

    The reference keeps its state in module globals, so one RefLib instance
    per variant library; `put`/`get` copy a whole MeshState in and out.
    """

    def __init__(self, variant: str):
        path = ref_path(variant)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        self.L = L
        self.variant = variant
        L.wxr_init.argtypes = [ctypes.c_int]
        L.wxr_consts.argtypes = [_dp]
        L.wxr_set_dxdy.argtypes = [_dp, _dp]
        L.wxr_put.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _ip, _ip, ctypes.c_int, ctypes.c_int]
        L.wxr_get.argtypes = [ctypes.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _ip]
        L.wxr_set_bounds.argtypes = [_ip, ctypes.c_int]
        L.wxr_set_jet.argtypes = [_dp]
        L.wxr_set_cooltab.argtypes = [_dp]
        L.wxr_cooldmc.argtypes = [ctypes.c_double]
        L.wxr_cooldmc.restype = ctypes.c_double
        L.wxr_uprim.argtypes = [_dp, _dp, _dp]
        L.wxr_primu.argtypes = [_dp, _dp]
        L.wxr_primf.argtypes = [_dp, _dp]
        L.wxr_limiter.argtypes = [_dp, _dp, _dp, _dp]
        L.wxr_fluxes.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.wxr_stepblock.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.wxr_viscosity.argtypes = [ctypes.c_int]
        L.wxr_boundary.argtypes = [ctypes.c_double, ctypes.c_int]
        L.wxr_atomic.argtypes = [ctypes.c_double, _dp]
        L.wxr_aloss.argtypes = [ctypes.c_double] * 6
        L.wxr_aloss.restype = ctypes.c_double
        L.wxr_cooling.argtypes = [ctypes.c_double]
        L.wxr_tstep.argtypes = [ctypes.c_double, ctypes.c_double]
        L.wxr_timestep.restype = ctypes.c_double
        L.wxr_criteria.argtypes = [ctypes.c_int, _ip, _ip]
        L.wxr_markref.argtypes = [_ip, _ip]
        L.wxr_critneighdown.argtypes = [_ip]
        L.wxr_refine.argtypes = [ctypes.c_int] * 5 + [_ip, _ip]
        L.wxr_coarse.argtypes = [ctypes.c_int] * 5
        L.wxr_locatebounds.argtypes = [ctypes.c_double, _ip, ctypes.c_int]
        L.wxr_locatebounds.restype = ctypes.c_int
        L.wxr_admesh.argtypes = []
        L.wxr_nbleafs.restype = ctypes.c_int
        L.wxr_output.argtypes = [ctypes.c_int]
        c = np.zeros(12)
        L.wxr_consts(dptr(c))
        self.consts = dict(cv=c[0], gamma=c[1], Tempsc=c[2], Psc=c[3], tsc=c[4], cfl=c[5],
                           eta=c[6], neq=int(c[7]), nlevs=int(c[8]), nblocks=int(c[9]),
                           rsc=c[10], vsc2=c[11])
        self.neq = self.consts["neq"]
        self.nblocks = self.consts["nblocks"]
        L.wxr_init(self.nblocks)

    def setup(self, params):
        dx = np.array(params.dx[:self.consts["nlevs"]])
        dy = np.array(params.dy[:self.consts["nlevs"]])
        self.L.wxr_set_dxdy(dptr(dx), dptr(dy))
        self.L.wxr_set_cooltab(dptr(np.array(params.cooltab[:])))
        self.L.wxr_set_jet(dptr(np.array(params.jet[:])))

    def put(self, m: MeshState):
        assert m.cap < self.nblocks and m.neq == self.neq
        self.L.wxr_put(m.cap, dptr(m.u), dptr(m.up), dptr(m.primit), iptr(m.lp),
                       iptr(m.leaf), iptr(m.icoord), m.s.nbmax, m.s.nbleafs)
        n = m.s.ninternal
        b = np.ascontiguousarray(m.bounds[:max(n, 1)])
        self.L.wxr_set_bounds(iptr(b), n)

    def get(self, m: MeshState):
        nbmax = np.zeros(1, dtype=np.int32)
        nbl = np.zeros(1, dtype=np.int32)
        self.L.wxr_get(m.cap, dptr(m.u), dptr(m.up), dptr(m.primit), iptr(m.lp),
                       iptr(m.leaf), iptr(m.icoord), iptr(nbmax), iptr(nbl))
        m.s.nbmax = int(nbmax[0])
        m.s.nbleafs = int(nbl[0])

    def flags(self):
        return (np.zeros(self.nblocks + 1, dtype=np.int32), np.zeros(self.nblocks + 1, dtype=np.int32))

    def locatebounds(self, time=0.0):
        """The reference's own locatebounds (src/locatebounds.f90:29-540) on the
        state last put: innerbounds(1:ninternal, 1:3) as an (n, 3) array."""
        cap = 4 * self.nblocks
        b = np.zeros((cap, 3), dtype=np.int32)
        n = self.L.wxr_locatebounds(float(time), iptr(b), cap)
        return b[:n].copy()

    def admesh(self):
        """The reference's own admesh (src/admesh.f90:4-111) on the state last put."""
        self.L.wxr_admesh()

    def uprim(self, uu):
        uu = np.ascontiguousarray(uu, dtype=np.float64)
        prim = np.zeros_like(uu)
        T = np.zeros(1)
        self.L.wxr_uprim(dptr(prim), dptr(uu), dptr(T))
        return prim, T[0]

    def primu(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        out = np.zeros_like(prim)
        self.L.wxr_primu(dptr(prim), dptr(out))
        return out

    def primf(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        out = np.zeros_like(prim)
        self.L.wxr_primf(dptr(prim), dptr(out))
        return out

    def limiter(self, pll, pl, pr, prr):
        pl = np.array(pl, dtype=np.float64)
        pr = np.array(pr, dtype=np.float64)
        self.L.wxr_limiter(dptr(np.ascontiguousarray(pll, dtype=np.float64)), dptr(pl), dptr(pr),
                           dptr(np.ascontiguousarray(prr, dtype=np.float64)))
        return pl, pr

    def fluxes(self, nb, choice):
        f = np.zeros((self.neq, NP, NP))
        g = np.zeros((self.neq, NP, NP))
        self.L.wxr_fluxes(nb, choice, dptr(f), dptr(g))
        return f, g

    def stepblock(self, nb, dt, f, g):
        self.L.wxr_stepblock(nb, dt, dptr(np.ascontiguousarray(f)), dptr(np.ascontiguousarray(g)))

    def atomic(self, dt, uu):
        uu = np.array(uu, dtype=np.float64)
        self.L.wxr_atomic(dt, dptr(uu))
        return uu

    def aloss(self, *a):
        return self.L.wxr_aloss(*[float(x) for x in a])

    def cooldmc(self, T):
        return self.L.wxr_cooldmc(float(T))
