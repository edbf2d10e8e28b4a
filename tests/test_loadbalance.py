"""Block placement plan of the MPI load balancer (src/loadbalance.f90):
Hilbert keys, the index sort and the new block IDs.

The oracle's restatement is pinned three ways: against the reference's own
hilbert2d and indexx, built from src/loadbalance.f90 into oracle/_ref/
(skipped where that build is absent); against the textbook Hilbert index
(the iterative rotate-and-reflect form of d(x, y)) and the curve's
properties; and, for the walk that assigns IDs (which lives in the MPI body
and is not built), against the properties that fix it: key order kept,
consecutive IDs per rank, nbleafs/nranks leaves per rank, each full rank
closing on a leaf.  The library's wx_lb_plan (host code, no device) must
equal the oracle exactly."""
import ctypes
import os

import numpy as np
import pytest

from walicxe2d_amd import capi
from walicxe2d_amd import params as P
from wxoracle import Oracle

from harness import oracle_mesh

REF_LB = os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref", "libwxref_lb.so")
_ip = ctypes.POINTER(ctypes.c_int32)


def _xy2d(n, x, y):
    """Textbook Hilbert index of (x, y) on an n x n grid (n a power of 2)."""
    d, s = 0, n // 2
    while s > 0:
        rx, ry = int(x & s > 0), int(y & s > 0)
        d += s * s * ((3 * rx) ^ ry)
        if ry == 0:
            if rx == 1:
                x, y = n - 1 - x, n - 1 - y
            x, y = y, x
        s //= 2
    return d


def test_hilbert_key_is_the_hilbert_curve():
    O = Oracle()
    for bits in range(1, 7):
        n = 1 << bits
        key = np.array([[O.hilbert2d(x, y, bits) for y in range(n)] for x in range(n)])
        assert sorted(key.ravel()) == list(range(n * n))          # a bijection onto 0..4^b-1
        pos = np.zeros((n * n, 2), dtype=int)
        for x in range(n):
            for y in range(n):
                pos[key[x, y]] = (x, y)
                assert key[x, y] == _xy2d(n, x, y)
                if bits > 1:                                       # the top digits are the parent's key
                    assert key[x, y] >> 2 == O.hilbert2d(x >> 1, y >> 1, bits - 1)
        assert np.all(np.abs(np.diff(pos, axis=0)).sum(axis=1) == 1)   # consecutive keys are neighbours
        assert key[0, 0] == 0
    # bits at or above `bits` do not enter the key
    assert O.hilbert2d(5 + 64, 3 + 128, 6) == O.hilbert2d(5, 3, 6)


def test_indexx_sorts_with_its_tie_order():
    O = Oracle()
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 6, 7, 11, 100, 2001):
        a = rng.integers(0, max(2, n // 4), n).astype(np.int32)
        ix = O.indexx(a)
        assert sorted(ix) == list(range(1, n + 1))
        assert np.all(np.diff(a[ix - 1]) >= 0)


@pytest.mark.skipif(not os.path.exists(REF_LB), reason="reference load balancer not built (make -C oracle ref)")
def test_oracle_matches_reference_routines():
    R = ctypes.CDLL(REF_LB)
    O = Oracle()
    rng = np.random.default_rng(1)
    for bits in (1, 3, 5, 7):
        x = rng.integers(0, 1 << (bits + 2), 500).astype(np.int32)      # includes bits above `bits`
        y = rng.integers(0, 1 << (bits + 2), 500).astype(np.int32)
        out = np.zeros(500, dtype=np.int32)
        R.wxr_lb_hilbert2d(x.ctypes.data_as(_ip), y.ctypes.data_as(_ip), out.ctypes.data_as(_ip), bits, 500)
        assert [O.hilbert2d(a, b, bits) for a, b in zip(x, y)] == list(out)
    for n in (1, 2, 4, 5, 6, 7, 8, 31, 500, 3000):
        for hi in (2, 8, 1 << 20):                                  # heavy ties .. nearly distinct
            a = rng.integers(0, hi, n).astype(np.int32)
            ref = np.zeros(n, dtype=np.int32)
            R.wxr_lb_indexx(n, a.ctypes.data_as(_ip), ref.ctypes.data_as(_ip))
            assert np.array_equal(O.indexx(a), ref), (n, hi)
    a = np.repeat(np.arange(40, dtype=np.int32), 7)[::-1].copy()     # sorted runs, reversed
    ref = np.zeros(len(a), dtype=np.int32)
    R.wxr_lb_indexx(len(a), a.ctypes.data_as(_ip), ref.ctypes.data_as(_ip))
    assert np.array_equal(O.indexx(a), ref)


def test_oracle_matches_reference_golden_vectors():
    """tests/golden/lb_vectors.npz: outputs of the reference's own hilbert2d
    and indexx on seeded inputs (tests/golden/make_lb_golden.py)."""
    O = Oracle()
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "lb_vectors.npz"))
    nh = ni = 0
    for name in g.files:
        if name.startswith("hilbert_") and name.endswith("_key"):
            bits = int(name.split("_")[1])
            x, y = g[f"hilbert_{bits}_x"], g[f"hilbert_{bits}_y"]
            assert [O.hilbert2d(a, b, bits) for a, b in zip(x, y)] == list(g[name]), bits
            nh += 1
        elif name.startswith("indexx_") and name.endswith("_idx"):
            a = g[name[:-4] + "_arr"]
            assert np.array_equal(O.indexx(a), g[name]), name
            ni += 1
    assert nh == 5 and ni == 21


def _jet_mesh(steps=30):
    """16x4 roots, 4 levels: the finest block index reaches 127, above the
    keys' 5 bits, so keys alias as in the reference and many tie (21 keys
    shared by up to 7 blocks after 30 steps), and indexx orders the ties
    differently from a stable sort."""
    cfg = P.config_c3(nbrootx=16, nbrooty=4, nblocks=6000)
    cfg.nlevs = 4
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 6000)
    for _ in range(4):
        O.admesh(pp, m)
    t = 0.0
    for _ in range(steps):
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        assert O.admesh(pp, m) == 0
    return cfg, O, m


def _check_walk(O, key, idnew, leaf, nbmax, nblocks, nranks):
    """The ID walk's defining properties, given the key order."""
    order = O.indexx(key[1:nbmax + 1])                              # block IDs in key order
    ids = idnew[order]
    rank = (ids - 1) // nblocks
    assert np.all(np.diff(rank) >= 0) and rank[0] == 0 and rank[-1] <= nranks - 1
    nleaf = int(np.count_nonzero(leaf[1:nbmax + 1]))
    quota = nleaf // nranks
    for r in range(rank[-1] + 1):
        sel = rank == r
        mine = ids[sel]
        assert np.array_equal(mine, r * nblocks + 1 + np.arange(len(mine)))
        lv = leaf[order[sel]] != 0
        if r < rank[-1]:                                            # a closed rank
            if quota > 0:
                assert lv.sum() == quota and lv[-1]                 # closes on its quota-th leaf
            else:
                assert len(mine) == 1                               # quota 0: one block each
    assert rank[-1] == nranks - 1 or nbmax < nranks                 # the last rank takes the rest


def test_plan_walk_properties_and_library_parity():
    cfg, O, m = _jet_mesh()
    nbmax = m.nbmax
    levels = {int(m.lp[nb, 0]) for nb in range(1, nbmax + 1)}
    assert len(levels) == 4
    _, key, _ = O.lb_plan(m.lp, m.leaf, m.icoord, nbmax, cfg.nlevs, cfg.nblocks, 1)
    _, cnt = np.unique(key[1:], return_counts=True)
    assert (cnt > 1).sum() >= 10                                    # ties to order
    assert not np.array_equal(O.indexx(key[1:]), np.argsort(key[1:], kind="stable") + 1)
    for nranks in (1, 2, 3, 4, 7, 8, 16):
        err, key, idnew = O.lb_plan(m.lp, m.leaf, m.icoord, nbmax, cfg.nlevs, cfg.nblocks, nranks)
        assert err == 0
        if nranks == 1:                                             # one rank: IDs in key order
            order = O.indexx(key[1:])
            assert np.array_equal(idnew[order], np.arange(1, nbmax + 1))
        _check_walk(O, key, idnew, m.leaf, nbmax, cfg.nblocks, nranks)
        k2, id2 = capi.lb_plan(m.lp, m.leaf, m.icoord, nbmax, cfg.nlevs, cfg.nblocks, nranks)
        assert np.array_equal(k2[1:], key[1:]) and np.array_equal(id2[1:], idnew[1:]), nranks
    # keys: the corner of each block on the finest grid
    err, key, _ = O.lb_plan(m.lp, m.leaf, m.icoord, nbmax, cfg.nlevs, cfg.nblocks, 1)
    for nb in range(1, nbmax + 1):
        sh = cfg.nlevs - int(m.lp[nb, 0])
        x, y = (int(m.icoord[nb, 0]) << sh) // 16, (int(m.icoord[nb, 1]) << sh) // 16
        assert key[nb] == _xy2d(1 << (cfg.nlevs + 1), x % (1 << (cfg.nlevs + 1)), y % (1 << (cfg.nlevs + 1)))


def test_plan_small_cases_and_errors():
    O = Oracle()
    # 2x1 roots, one level: keys of blocks (0,0) and (1,0) over 2 bits
    lp = np.zeros((3, 8), dtype=np.int32)
    lp[1:, 0] = 1
    leaf = np.array([0, 1, 1], dtype=np.int32)
    ic = np.array([[0, 0], [0, 0], [16, 0]], dtype=np.int32)
    err, key, idnew = O.lb_plan(lp, leaf, ic, 2, 1, 10, 2)
    assert err == 0 and list(key[1:]) == [_xy2d(4, 0, 0), _xy2d(4, 1, 0)]
    assert list(idnew[1:]) == [1, 11]                                # one leaf each, rank 1 starts at 11
    k2, id2 = capi.lb_plan(lp, leaf, ic, 2, 1, 10, 2)
    assert list(k2[1:]) == list(key[1:]) and list(id2[1:]) == [1, 11]
    # more ranks than leaves: quota 0, one block per rank until the last
    err, _, idnew = O.lb_plan(lp, leaf, ic, 2, 1, 10, 5)
    assert err == 0 and list(idnew[1:]) == [1, 11]
    # capacity reached: the reference stops, the library reports WX_EBLOCKS
    cfg, _, m = _jet_mesh(steps=0)
    err, _, _ = O.lb_plan(m.lp, m.leaf, m.icoord, m.nbmax, cfg.nlevs, 8, 1)
    assert err == 2
    with pytest.raises(capi.WxError, match="status 2"):
        capi.lb_plan(m.lp, m.leaf, m.icoord, m.nbmax, cfg.nlevs, 8, 1)
    with pytest.raises(capi.WxError, match="status 3"):
        capi.lb_plan(m.lp, m.leaf, m.icoord, m.nbmax, cfg.nlevs, 3000, 0)


@pytest.mark.gpu
def test_plan_of_device_tree_matches_oracle():
    """The plan of the tree the device's own admesh built (wx_get_tree) equals
    the oracle's plan of its mesh after the same steps."""
    from walicxe2d_amd.capi import Simulation
    cfg = P.config_c3(nbrootx=16, nbrooty=4, nblocks=6000)
    cfg.nlevs = 4
    O = Oracle()
    pp, m = oracle_mesh(O, cfg, 6000)
    sim = Simulation(cfg, nblocks=6000)
    sim.basegrid()
    for _ in range(4):
        O.admesh(pp, m)
        sim.admesh()
    t = 0.0
    for _ in range(30):
        dt = O.timestep(pp, m)
        t += dt
        assert O.tstep(pp, m, t, dt) == 0
        sim.step(1, admesh=False)
        assert O.admesh(pp, m) == 0
        sim.admesh()
    lp, leaf, ic, nbmax, _ = sim.tree()
    sim.close()
    assert nbmax == m.nbmax
    for nranks in (1, 4, 8):
        err, key, idnew = O.lb_plan(m.lp, m.leaf, m.icoord, nbmax, cfg.nlevs, cfg.nblocks, nranks)
        k2, id2 = capi.lb_plan(lp, leaf, ic, nbmax, cfg.nlevs, cfg.nblocks, nranks)
        assert err == 0 and np.array_equal(k2[1:], key[1:]) and np.array_equal(id2[1:], idnew[1:])
