"""The oracle's mesh layer against the reference's own Fortran (oracle/_ref),
bit for bit: the ghost-fill list of locatebounds in its order (the order
decides which block the OUTFYB branch writes, src/boundaryI.f90:87, and
which write of a ghost cell is the last), the flags of markref + critneighup
and critneighdown, the block IDs, pointers and data of admesh
(src/admesh.f90:4-111, incl. the reuse of stale son IDs), and whole runs of
the time loop (src/main.f90:87-98: timestep, tstep with the reference's own
locatebounds, admesh) -- the full 200-step C1 run among them.  Skips where
the reference tree is absent (oracle/_ref not built)."""
import os
import sys

import numpy as np
import pytest

from wxoracle import Oracle, RefLib, ref_available

from harness import oracle_mesh

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from mesh_inputs import LB_CASES, LOOP_CASES, cfg_for, random_mesh  # noqa: E402

O = Oracle()
_libs = {}


def ref(variant):
    if not ref_available(variant):
        pytest.skip(f"oracle/_ref/libwxref_{variant}.so not built (reference tree absent)")
    if variant not in _libs:
        _libs[variant] = RefLib(variant)
    return _libs[variant]


def same_tree(a, b):
    n = a.nbmax
    return (a.nbmax == b.nbmax and a.nbleafs == b.nbleafs
            and np.array_equal(a.lp[1:n + 1], b.lp[1:n + 1])
            and np.array_equal(a.leaf[1:n + 1], b.leaf[1:n + 1])
            and np.array_equal(a.icoord[1:n + 1], b.icoord[1:n + 1]))


@pytest.mark.parametrize("variant,seed", LB_CASES)
def test_locatebounds_order_bitexact(variant, seed):
    """innerbounds(1:ninternal, 1:3) of the reference's locatebounds equals
    the oracle's list entry for entry, on random multi-level meshes (box
    faces, same-level and coarse/fine entries in every direction)."""
    R = ref(variant)
    cfg, pp, m = random_mesh(O, variant, seed)
    R.setup(pp)
    n = O.locatebounds(pp, m)
    R.put(m)
    b = R.locatebounds(0.0)
    assert b.shape[0] == n, (b.shape, n)
    assert np.array_equal(b, m.bounds[:n])
    if cfg.nlevs > 1:
        assert (b[:, 2] >= 1).any() and np.isin(b[:, 2], [2, 3, 5, 6, 8, 9, 11, 12]).any()


@pytest.mark.parametrize("variant", ["c2", "c3", "hll"])
def test_markref_critneigh_bitexact(variant):
    """markref + critneighup (src/markref.f90, src/critneighup.f90) and
    critneighdown (src/critneighdown.f90) flags by block ID."""
    R = ref(variant)
    for seed in range(4):
        cfg, pp, m = random_mesh(O, variant, 10 + seed)
        R.setup(pp)
        R.put(m)
        up_o, dn_o = O.markref(pp, m)
        O.critneighup(pp, m, up_o)
        up_r, dn_r = R.flags()
        R.L.wxr_markref(up_r.ctypes.data_as(type(m.s.lp)), dn_r.ctypes.data_as(type(m.s.lp)))
        assert np.array_equal(up_o, up_r[:m.cap + 1]) and np.array_equal(dn_o, dn_r[:m.cap + 1])
        # critneighdown on random coarsening flags over the leaves
        rng = np.random.default_rng(seed)
        dn = np.zeros(m.cap + 1, dtype=np.int32)
        dn[m.leaves()] = rng.integers(0, 2, len(m.leaves()))
        dn_r = np.zeros(R.nblocks + 1, dtype=np.int32)
        dn_r[:m.cap + 1] = dn
        O.critneighdown(pp, m, dn)
        R.L.wxr_critneighdown(dn_r.ctypes.data_as(type(m.s.lp)))
        assert np.array_equal(dn, dn_r[:m.cap + 1])


@pytest.mark.parametrize("variant", ["c2", "c3", "hll"])
def test_admesh_bitexact(variant):
    """One admesh from the same state: block IDs (new sons at nbmax + 1..4,
    stale son IDs reused), pointers, leaf flags, coordinates and the
    refined / coarsened data of every block."""
    R = ref(variant)
    for seed in range(3):
        cfg, pp, m = random_mesh(O, variant, 20 + seed)
        R.setup(pp)
        R.put(m)
        assert O.admesh(pp, m) == 0
        R.admesh()
        m2 = m.copy()
        R.get(m2)
        assert same_tree(m, m2)
        ids = list(range(1, m.nbmax + 1))
        assert np.array_equal(m.u[ids], m2.u[ids])
        assert np.array_equal(m.primit[m.leaves()], m2.primit[m.leaves()])


@pytest.mark.parametrize("variant,steps", LOOP_CASES)
def test_time_loop_bitexact(variant, steps):
    """The reference's time loop on its own (timestep, time += dt, tstep with
    its own locatebounds, admesh; after the initial admesh passes of
    src/main.f90:60-63) against the oracle's from the same initial state:
    dt, every cell of every block, the primitives and the tree equal at every
    step.  C1 runs its full 200 steps."""
    R = ref(variant)
    cfg = cfg_for(variant)
    pp, m = oracle_mesh(O, cfg, 3000)
    R.setup(pp)
    R.put(m)
    amr = cfg.nlevs > 1
    if amr:
        for _ in range(3):
            assert O.admesh(pp, m) == 0
            R.admesh()
    t_o = t_r = 0.0
    m2 = m.copy()
    for s in range(steps):
        dt_o = O.timestep(pp, m)
        dt_r = R.L.wxr_timestep()
        assert dt_o == dt_r, (s, dt_o, dt_r)
        t_o += dt_o
        t_r += dt_r
        assert O.tstep(pp, m, t_o, dt_o) == 0
        R.L.wxr_tstep(t_r, dt_r)
        if amr:
            assert O.admesh(pp, m) == 0
            R.admesh()
        R.get(m2)
        assert same_tree(m, m2), (variant, s)
        lv = m.leaves()
        assert np.array_equal(m.u[lv], m2.u[lv]), (variant, s)
        assert np.array_equal(m.primit[lv], m2.primit[lv]), (variant, s)
