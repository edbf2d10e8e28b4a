"""ctypes binding of the C-ABI (include/wx_capi.h) and a host-side driver.

`Simulation` mirrors the reference's driver-level interface (src/main.f90
time loop: timestep -> time += dt -> tstep -> admesh) on one GPU-resident AMR
mesh.  It refuses to run without the HIP library: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import NCELL, NP, Config, WxParams, initial_state

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libwxhip.so")

WX_LAYOUT_BLOCK = 0
WX_LAYOUT_FORTRAN = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lib = None

EXPORTS = [
    "wx_create", "wx_destroy", "wx_upload", "wx_download", "wx_get_tree", "wx_set_bounds",
    "wx_locatebounds", "wx_timestep", "wx_tstep", "wx_criteria", "wx_refine", "wx_coarse",
    "wx_upload_tree", "wx_basegrid", "wx_admesh", "wx_step", "wx_set_time", "wx_info",
    "wx_last_error", "wx_stream", "wx_profile",
    "wx_ens_create", "wx_ens_destroy", "wx_ens_basegrid", "wx_ens_admesh", "wx_ens_step", "wx_ens_instance",
    "wx_ens_download", "wx_ens_get_tree", "wx_ens_set_graph", "wx_ens_set_launch", "wx_ens_tune",
    "wx_ens_profile", "wx_ens_info", "wx_ens_create_groups", "wx_ens_groups",
    "wx_shape", "wx_write_bin", "wx_read_bin", "wx_ens_shape", "wx_ens_write_bin",
    "wx_sur_pairs", "wx_sur_features", "wx_sur_infer", "wx_sur_grad", "wx_sur_train", "wx_set_surrogate",
    "wx_ens_set_surrogate", "wx_sur_counters", "wx_ens_sur_counters", "wx_lb_plan", "wx_sur_infer_noalloc",
    "wx_set_builder", "wx_get_tables",
]

_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


class WxSurModel(ctypes.Structure):
    """include/wx_capi.h wx_sur_model."""
    _fields_ = [("theta", ctypes.c_float * 8770), ("in_mu", ctypes.c_float * 4), ("in_sig", ctypes.c_float * 4),
                ("out_mu", ctypes.c_float * 2), ("out_sig", ctypes.c_float * 2), ("lo", ctypes.c_float * 4),
                ("hi", ctypes.c_float * 4), ("tol", ctypes.c_float), ("pad", ctypes.c_int32 * 3)]

KERNELS = ["stageA", "stageB", "fillI_cool", "dtreduce", "blockstat", "meshops", "host_admesh", "host_tables",
           "builder"]


class WxError(RuntimeError):
    pass


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Loads libwxhip.so (the module's LIBPATH unless a path is given); raises
    if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIBPATH
    if not os.path.exists(path):
        raise WxError(f"{path} is missing: build it with `python -m walicxe2d_amd.build` "
                      "(the hot path has no CPU fallback)")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.wx_create.argtypes = [ctypes.POINTER(WxParams), ctypes.POINTER(P)]
    L.wx_destroy.argtypes = [P]
    L.wx_upload.argtypes = [P, _dp, ctypes.c_int32, _ip, _ip, _ip, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.wx_download.argtypes = [P, _dp, _dp, ctypes.c_int32, ctypes.c_int32]
    L.wx_get_tree.argtypes = [P, _ip, _ip, _ip, ctypes.c_int32, _ip, _ip]
    L.wx_set_bounds.argtypes = [P, _ip, ctypes.c_int32, ctypes.c_int32]
    L.wx_locatebounds.argtypes = [P, _ip]
    L.wx_set_builder.argtypes = [P, ctypes.c_int32]
    L.wx_get_tables.argtypes = [P, _ip, _ip, _ip, _ip, ctypes.c_int32]
    L.wx_timestep.argtypes = [P, _dp]
    L.wx_tstep.argtypes = [P, ctypes.c_double, ctypes.c_double]
    L.wx_criteria.argtypes = [P, _ip, _ip]
    L.wx_refine.argtypes = [P] + [ctypes.c_int32] * 5
    L.wx_coarse.argtypes = [P] + [ctypes.c_int32] * 5
    L.wx_upload_tree.argtypes = [P, _ip, _ip, _ip, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.wx_basegrid.argtypes = [P, _dp, ctypes.c_int32]
    L.wx_admesh.argtypes = [P, _ip]
    L.wx_step.argtypes = [P, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp]
    L.wx_set_time.argtypes = [P, ctypes.c_double]
    L.wx_info.argtypes = [P, _ip, _ip, _ip]
    L.wx_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int32]
    L.wx_stream.argtypes = [P]
    L.wx_lb_plan.argtypes = [_ip, _ip, _ip] + [ctypes.c_int32] * 5 + [_ip, _ip]
    L.wx_stream.restype = P
    L.wx_profile.argtypes = [P, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    L.wx_ens_create.argtypes = [ctypes.POINTER(WxParams), ctypes.c_int32, ctypes.POINTER(P)]
    L.wx_ens_create_groups.argtypes = [ctypes.POINTER(WxParams), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(P)]
    L.wx_ens_groups.argtypes = [P, _ip]
    L.wx_shape.argtypes = [P, _ip, _ip, _ip]
    L.wx_write_bin.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    L.wx_read_bin.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p)]
    L.wx_ens_shape.argtypes = [P, ctypes.c_int32, _ip, _ip, _ip]
    L.wx_ens_write_bin.argtypes = [P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    L.wx_ens_destroy.argtypes = [P]
    L.wx_ens_basegrid.argtypes = [P, ctypes.c_int32, _dp, ctypes.c_int32]
    L.wx_ens_admesh.argtypes = [P, _ip]
    L.wx_ens_step.argtypes = [P, ctypes.c_int32, ctypes.c_int32, _dp]
    L.wx_ens_instance.argtypes = [P, ctypes.c_int32, _dp, _dp, _ip, _ip]
    L.wx_ens_download.argtypes = [P, ctypes.c_int32, _dp]
    L.wx_ens_get_tree.argtypes = [P, ctypes.c_int32, _ip, _ip, _ip, ctypes.c_int32, _ip, _ip]
    L.wx_ens_set_graph.argtypes = [P, ctypes.c_int32]
    L.wx_ens_set_launch.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.wx_ens_tune.argtypes = [P, ctypes.c_int32, _ip, _ip, _dp]
    L.wx_ens_profile.argtypes = [P, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    L.wx_ens_info.argtypes = [P, _ip, _ip, _ip, _ip]
    SM = ctypes.POINTER(WxSurModel)
    L.wx_sur_pairs.argtypes = [ctypes.POINTER(WxParams), _dp, ctypes.c_int64, _dp]
    L.wx_sur_features.argtypes = [ctypes.POINTER(WxParams), _dp, ctypes.c_int64, _fp, _fp]
    L.wx_sur_infer.argtypes = [SM, _fp, ctypes.c_int64, _fp]
    L.wx_sur_infer_noalloc.argtypes = [P, P, ctypes.c_int64, P, P]
    L.wx_sur_grad.argtypes = [SM, _fp, _fp, ctypes.c_int64, ctypes.c_float, ctypes.c_float, _fp, _fp]
    L.wx_sur_train.argtypes = [SM, _fp, _fp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_float, _fp]
    L.wx_set_surrogate.argtypes = [P, SM]
    L.wx_ens_set_surrogate.argtypes = [P, SM]
    L.wx_sur_counters.argtypes = [P, ctypes.c_int32, _i64p, _i64p]
    L.wx_ens_sur_counters.argtypes = [P, ctypes.c_int32, _i64p, _i64p]
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(512)
    load_library().wx_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise WxError(f"{what} failed (status {rc}): {last_error()}")


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def tree_to_fortran(lp: np.ndarray, leaf: np.ndarray, icoord: np.ndarray, nbmax: int):
    """Row-per-ID arrays (index 0 unused) -> Fortran column-major lp(ld,8),
    leaf(ld), icoord(ld,2) with ld = nbmax."""
    ld = max(nbmax, 1)
    flp = np.ascontiguousarray(lp[1:ld + 1].T, dtype=np.int32)
    fleaf = np.ascontiguousarray(leaf[1:ld + 1], dtype=np.int32)
    fic = np.ascontiguousarray(icoord[1:ld + 1].T, dtype=np.int32)
    return flp, fleaf, fic, ld


def lb_plan(lp: np.ndarray, leaf: np.ndarray, icoord: np.ndarray, nbmax: int, nlevs: int, nblocks: int,
            nranks: int):
    """wx_lb_plan on row-per-ID tree arrays (index 0 unused, as tree()
    returns them): (hilbert[nbmax+1], idnew[nbmax+1]), slot 0 unused."""
    L = load_library()
    flp, fleaf, fic, ld = tree_to_fortran(lp, leaf, icoord, nbmax)
    key = np.zeros(nbmax + 1, dtype=np.int32)
    idnew = np.zeros(nbmax + 1, dtype=np.int32)
    _check(L.wx_lb_plan(flp.ctypes.data_as(_ip), fleaf.ctypes.data_as(_ip), fic.ctypes.data_as(_ip), ld,
                        nbmax, nlevs, nblocks, nranks, key[1:].ctypes.data_as(_ip),
                        idnew[1:].ctypes.data_as(_ip)), "wx_lb_plan")
    return key, idnew


class Simulation:
    """One AMR mesh resident on the GPU behind the C-ABI."""

    def __init__(self, cfg: Config, nblocks: int | None = None):
        self.L = load_library()
        self.cfg = cfg
        if nblocks is not None:
            cfg.nblocks = nblocks
        self.params = cfg.params()
        self.neq = cfg.neq
        self.cap = cfg.nblocks
        h = ctypes.c_void_p()
        _check(self.L.wx_create(ctypes.byref(self.params), ctypes.byref(h)), "wx_create")
        self.h = h
        self.time = 0.0

    def close(self):
        if getattr(self, "h", None):
            self.L.wx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up -----------------------------------------------------
    def basegrid(self):
        """Root layout + initial conditions (src/basegrid.f90:26-97 generalised)."""
        u, lp, icoord = initial_state(self.cfg)
        nroots = lp.shape[0] - 1
        _check(self.L.wx_basegrid(self.h, _d(np.ascontiguousarray(u)), nroots), "wx_basegrid")
        self.time = 0.0
        _check(self.L.wx_set_time(self.h, 0.0), "wx_set_time")

    def upload(self, u: np.ndarray, lp: np.ndarray, leaf: np.ndarray, icoord: np.ndarray,
               nbmax: int, nbleafs: int):
        """u: [>= nbmax+1][neq][20][20] block layout; tree arrays row-per-ID."""
        flp, fleaf, fic, ld = tree_to_fortran(lp, leaf, icoord, nbmax)
        u = np.ascontiguousarray(u[:nbmax + 1], dtype=np.float64)
        _check(self.L.wx_upload(self.h, _d(u), WX_LAYOUT_BLOCK, _i(flp), _i(fleaf), _i(fic), ld,
                                nbmax, nbleafs), "wx_upload")

    def upload_tree(self, lp, leaf, icoord, nbmax, nbleafs):
        flp, fleaf, fic, ld = tree_to_fortran(lp, leaf, icoord, nbmax)
        _check(self.L.wx_upload_tree(self.h, _i(flp), _i(fleaf), _i(fic), ld, nbmax, nbleafs),
               "wx_upload_tree")

    # ---- state --------------------------------------------------------
    def info(self):
        a, b, c = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32())
        _check(self.L.wx_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "wx_info")
        return a.value, b.value, c.value

    def download(self, prim: bool = False):
        nbmax = self.info()[0]
        u = np.zeros((nbmax + 1, self.neq, NP, NP))
        pr = np.zeros_like(u) if prim else None
        _check(self.L.wx_download(self.h, _d(u), _d(pr) if prim else None, WX_LAYOUT_BLOCK, 0), "wx_download")
        return (u, pr) if prim else u

    def download_fortran(self, ld: int = 0):
        nbmax = self.info()[0]
        ld = ld or nbmax
        u = np.zeros(ld * self.neq * NCELL)
        _check(self.L.wx_download(self.h, _d(u), None, WX_LAYOUT_FORTRAN, ld), "wx_download")
        return u

    def tree(self):
        nbmax = self.info()[0]
        ld = max(nbmax, 1)
        flp = np.zeros((8, ld), dtype=np.int32)
        fleaf = np.zeros(ld, dtype=np.int32)
        fic = np.zeros((2, ld), dtype=np.int32)
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L.wx_get_tree(self.h, _i(flp), _i(fleaf), _i(fic), ld, ctypes.byref(a), ctypes.byref(b)),
               "wx_get_tree")
        lp = np.zeros((ld + 1, 8), dtype=np.int32)
        leaf = np.zeros(ld + 1, dtype=np.int32)
        icoord = np.zeros((ld + 1, 2), dtype=np.int32)
        lp[1:] = flp.T
        leaf[1:] = fleaf
        icoord[1:] = fic.T
        return lp, leaf, icoord, a.value, b.value

    # ---- the reference's interfaces ----------------------------------
    def set_bounds(self, bounds: np.ndarray):
        """innerbounds rows (n,3) in list order."""
        b = np.asarray(bounds, dtype=np.int32).reshape(-1, 3)
        n = b.shape[0]
        fb = np.ascontiguousarray(b.T) if n else np.zeros((3, 1), dtype=np.int32)
        _check(self.L.wx_set_bounds(self.h, _i(fb), max(n, 1), n), "wx_set_bounds")

    def locatebounds(self) -> int:
        n = ctypes.c_int32()
        _check(self.L.wx_locatebounds(self.h, ctypes.byref(n)), "wx_locatebounds")
        return n.value

    def set_builder(self, on_device):
        """Build the gather ops on the device (True / 1) or on the host
        (False / 0); 2: on the device with the many-workgroup scan at any
        size; 3 (the default): on the device from 3072 leaves up."""
        mode = on_device if isinstance(on_device, int) and not isinstance(on_device, bool) else (1 if on_device else 0)
        _check(self.L.wx_set_builder(self.h, mode), "wx_set_builder")

    def tables(self):
        """The current gather tables: leaf records (nleaf, 8), ops of fills II
        and I (nleaf, 8, 2) as int32 arrays."""
        n = ctypes.c_int32()
        _check(self.L.wx_get_tables(self.h, ctypes.byref(n), None, None, None, 0), "wx_get_tables")
        m = n.value
        rec = np.zeros((m, 8), dtype=np.int32)
        o2 = np.zeros((m, 8, 2), dtype=np.int32)
        o1 = np.zeros((m, 8, 2), dtype=np.int32)
        _check(self.L.wx_get_tables(self.h, ctypes.byref(n), _i(rec), _i(o2), _i(o1), m), "wx_get_tables")
        return rec, o2, o1

    def timestep(self) -> float:
        dt = np.zeros(1)
        _check(self.L.wx_timestep(self.h, _d(dt)), "wx_timestep")
        return float(dt[0])

    def tstep(self, time: float, dt: float):
        _check(self.L.wx_tstep(self.h, time, dt), "wx_tstep")

    def criteria(self):
        up = np.zeros(self.cap, dtype=np.int32)
        dn = np.zeros(self.cap, dtype=np.int32)
        _check(self.L.wx_criteria(self.h, _i(up), _i(dn)), "wx_criteria")
        return up, dn

    def refine(self, nb, s1, s2, s3, s4):
        _check(self.L.wx_refine(self.h, nb, s1, s2, s3, s4), "wx_refine")

    def coarse(self, dad, s1, s2, s3, s4):
        _check(self.L.wx_coarse(self.h, dad, s1, s2, s3, s4), "wx_coarse")

    def admesh(self) -> int:
        ch = ctypes.c_int32()
        _check(self.L.wx_admesh(self.h, ctypes.byref(ch)), "wx_admesh")
        return ch.value

    def profile(self, enable: int = -1):
        """Per-kernel (ms, launches) since the last reset; enable 1/0 resets and
        switches event recording on/off, -1 only reads."""
        ms = np.zeros(len(KERNELS))
        n = np.zeros(len(KERNELS), dtype=np.int64)
        _check(self.L.wx_profile(self.h, enable, _d(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 len(KERNELS)),
               "wx_profile")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS)}

    def set_time(self, t: float):
        """Simulation time (e.g. itprint * dtprint after a warm start)."""
        _check(self.L.wx_set_time(self.h, float(t)), "wx_set_time")
        self.time = float(t)

    def write_bin(self, pointers_path: str, blocks_path: str, logical_true: int = -1):
        """The reference's OUTBIN checkpoint pair (src/output.f90:39-67)."""
        _check(self.L.wx_write_bin(self.h, os.fsencode(pointers_path), os.fsencode(blocks_path), logical_true),
               "wx_write_bin")

    def read_bin(self, pointers_paths, blocks_paths):
        """Warm start from one pair per rank (src/basegrid.f90:167-251)."""
        n = len(pointers_paths)
        if n != len(blocks_paths) or n < 1:
            raise WxError("one blocks file per pointers file")
        pp = (ctypes.c_char_p * n)(*[os.fsencode(x) for x in pointers_paths])
        bp = (ctypes.c_char_p * n)(*[os.fsencode(x) for x in blocks_paths])
        _check(self.L.wx_read_bin(self.h, n, pp, bp), "wx_read_bin")

    def set_surrogate(self, model):
        """Cooling surrogate (a walicxe2d_amd.surrogate.SurrogateModel) in the
        coolingh pass; None restores the exact routine everywhere."""
        _check(self.L.wx_set_surrogate(self.h, ctypes.byref(model.to_c()) if model is not None else None),
               "wx_set_surrogate")

    def sur_counters(self, reset: bool = False):
        """(cells the surrogate pass handled, cells sent to the exact routine)."""
        c, f = ctypes.c_int64(), ctypes.c_int64()
        _check(self.L.wx_sur_counters(self.h, 1 if reset else 0, ctypes.byref(c), ctypes.byref(f)),
               "wx_sur_counters")
        return c.value, f.value

    def step(self, nsteps: int = 1, admesh: bool = True):
        """nsteps iterations of src/main.f90:87-98; returns (time, dt, zone_updates)."""
        t, dt, z = np.zeros(1), np.zeros(1), np.zeros(1)
        _check(self.L.wx_step(self.h, nsteps, 1 if admesh else 0, _d(t), _d(dt), _d(z)), "wx_step")
        self.time = float(t[0])
        return float(t[0]), float(dt[0]), float(z[0])


class Ensemble:
    """Many independent AMR meshes (one Config each, same neq) advanced
    together on the GPU behind the wx_ens_* entry points: each stage is one
    launch over all instances' leaf blocks of a group; `groups` groups of
    consecutive instances are stepped concurrently, each on its own stream."""

    def __init__(self, cfgs, groups: int = 1):
        self.L = load_library()
        self.cfgs = list(cfgs)
        self.n = len(self.cfgs)
        if self.n < 1:
            raise WxError("an ensemble needs at least one instance")
        self.neq = self.cfgs[0].neq
        arr = (WxParams * self.n)(*[c.params() for c in self.cfgs])
        h = ctypes.c_void_p()
        _check(self.L.wx_ens_create_groups(arr, self.n, groups, ctypes.byref(h)), "wx_ens_create_groups")
        self.h = h
        self.groups = groups

    def close(self):
        if getattr(self, "h", None):
            self.L.wx_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def basegrid(self):
        for i, c in enumerate(self.cfgs):
            u, lp, _ = initial_state(c)
            _check(self.L.wx_ens_basegrid(self.h, i, _d(np.ascontiguousarray(u)), lp.shape[0] - 1),
                   "wx_ens_basegrid")

    def admesh(self) -> int:
        ch = ctypes.c_int32()
        _check(self.L.wx_ens_admesh(self.h, ctypes.byref(ch)), "wx_ens_admesh")
        return ch.value

    def step(self, nsteps: int = 1, admesh: bool = True) -> float:
        """nsteps iterations for every instance; returns the zone-updates."""
        z = np.zeros(1)
        _check(self.L.wx_ens_step(self.h, nsteps, 1 if admesh else 0, _d(z)), "wx_ens_step")
        return float(z[0])

    def instance(self, i: int):
        """(time, dt, nbmax, nbleafs) of instance i."""
        t, dt = np.zeros(1), np.zeros(1)
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L.wx_ens_instance(self.h, i, _d(t), _d(dt), ctypes.byref(a), ctypes.byref(b)),
               "wx_ens_instance")
        return float(t[0]), float(dt[0]), a.value, b.value

    def download(self, i: int):
        nbmax = self.instance(i)[2]
        u = np.zeros((nbmax + 1, self.neq, NP, NP))
        _check(self.L.wx_ens_download(self.h, i, _d(u)), "wx_ens_download")
        return u

    def tree(self, i: int):
        nbmax = self.instance(i)[2]
        ld = max(nbmax, 1)
        flp = np.zeros((8, ld), dtype=np.int32)
        fleaf = np.zeros(ld, dtype=np.int32)
        fic = np.zeros((2, ld), dtype=np.int32)
        a, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self.L.wx_ens_get_tree(self.h, i, _i(flp), _i(fleaf), _i(fic), ld, ctypes.byref(a),
                                      ctypes.byref(b)), "wx_ens_get_tree")
        lp = np.zeros((ld + 1, 8), dtype=np.int32)
        leaf = np.zeros(ld + 1, dtype=np.int32)
        icoord = np.zeros((ld + 1, 2), dtype=np.int32)
        lp[1:] = flp.T
        leaf[1:] = fleaf
        icoord[1:] = fic.T
        return lp, leaf, icoord, a.value, b.value

    def set_graph(self, on: bool):
        _check(self.L.wx_ens_set_graph(self.h, 1 if on else 0), "wx_ens_set_graph")

    def set_launch(self, kernel: int, grid: int, threads: int = 256):
        """Launch shape of stage kernel `kernel` (0 A, 1 B, 2 fill I + cooling):
        persistent workgroups (0 = one per leaf block) and threads per workgroup."""
        _check(self.L.wx_ens_set_launch(self.h, kernel, grid, threads), "wx_ens_set_launch")

    def tune(self, reps: int = 5):
        """Autotunes the stage kernels' launch shapes; returns (grids, threads, us per launch)."""
        g = np.zeros(3, dtype=np.int32)
        t = np.zeros(3, dtype=np.int32)
        us = np.zeros(3)
        _check(self.L.wx_ens_tune(self.h, reps, _i(g), _i(t), _d(us)), "wx_ens_tune")
        return [int(x) for x in g], [int(x) for x in t], [float(x) for x in us]

    def info(self):
        """(instances, leaf blocks in the queue, stage grids, stage threads)."""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        g = np.zeros(3, dtype=np.int32)
        t = np.zeros(3, dtype=np.int32)
        _check(self.L.wx_ens_info(self.h, ctypes.byref(a), ctypes.byref(b), _i(g), _i(t)), "wx_ens_info")
        return a.value, b.value, [int(x) for x in g], [int(x) for x in t]

    def write_bin(self, i: int, pointers_path: str, blocks_path: str, logical_true: int = -1):
        """The reference's OUTBIN checkpoint pair of instance i."""
        _check(self.L.wx_ens_write_bin(self.h, i, os.fsencode(pointers_path), os.fsencode(blocks_path),
                                       logical_true), "wx_ens_write_bin")

    def set_surrogate(self, model):
        _check(self.L.wx_ens_set_surrogate(self.h, ctypes.byref(model.to_c()) if model is not None else None),
               "wx_ens_set_surrogate")

    def sur_counters(self, reset: bool = False):
        c, f = ctypes.c_int64(), ctypes.c_int64()
        _check(self.L.wx_ens_sur_counters(self.h, 1 if reset else 0, ctypes.byref(c), ctypes.byref(f)),
               "wx_ens_sur_counters")
        return c.value, f.value

    def profile(self, enable: int = -1):
        ms = np.zeros(len(KERNELS))
        n = np.zeros(len(KERNELS), dtype=np.int64)
        _check(self.L.wx_ens_profile(self.h, enable, _d(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     len(KERNELS)),
               "wx_ens_profile")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNELS)}
