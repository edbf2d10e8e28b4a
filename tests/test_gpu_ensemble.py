"""Batched ensemble (wx_ens_* through the C-ABI) against single-mesh runs of
the same instances (wx_ctx): states, trees, times and timesteps must be
bit-identical, since both run the same kernels on the same data.  Also:
hipGraph replays against eager launches, and autotuned launch shapes against
the default ones, step by step."""
import numpy as np
import pytest

from walicxe2d_amd import params as P
from walicxe2d_amd.capi import Ensemble, Simulation

pytestmark = pytest.mark.gpu

# explosion-ensemble instances with max levels 3, 3, 4, 5, 6 (seeded draws)
C4_PICK = [0, 2, 9, 8, 7]


def _c4(i):
    return P.config_c4(i, nblocks=3000)


def _leaf_tiles(u, leaf):
    ids = np.nonzero(leaf)[0]
    return ids, u[ids]


def _compare_instance(ens, i, sim, label):
    t_e, dt_e, nbmax_e, nl_e = ens.instance(i)
    lp_s, leaf_s, ic_s, nbmax_s, nl_s = sim.tree()
    lp_e, leaf_e, ic_e, _, _ = ens.tree(i)
    assert (nbmax_e, nl_e) == (nbmax_s, nl_s), f"{label}: nbmax/leaves {nbmax_e}/{nl_e} vs {nbmax_s}/{nl_s}"
    assert np.array_equal(lp_e, lp_s) and np.array_equal(leaf_e, leaf_s) and np.array_equal(ic_e, ic_s), label
    assert t_e == sim.time, f"{label}: time {t_e!r} vs {sim.time!r}"
    ids, a = _leaf_tiles(ens.download(i), leaf_e)
    _, b = _leaf_tiles(sim.download(), leaf_s)
    same = np.array_equal(a, b)
    if not same:
        d = np.abs(a - b)
        k = np.unravel_index(int(np.argmax(d)), d.shape)
        raise AssertionError(f"{label}: tiles differ, max |diff| {d.max():.3e} at leaf {ids[k[0]]} var {k[1]} "
                             f"cell ({k[3]},{k[2]})")
    return dt_e


def _run_pair(cfgs, nsteps, label, groups=1, graph=False):
    ens = Ensemble(cfgs, groups=groups)
    ens.basegrid()
    sims = []
    for c in cfgs:
        s = Simulation(c)
        s.basegrid()
        sims.append(s)
    amr = any(c.nlevs > 1 for c in cfgs)
    if amr:
        for _ in range(3):
            ens.admesh()
            for s in sims:
                s.admesh()
    ens.set_graph(graph)
    for step in range(nsteps):
        ens.step(1, admesh=amr)
        for s, c in zip(sims, cfgs):
            s.step(1, admesh=c.nlevs > 1)
        for i, s in enumerate(sims):
            _compare_instance(ens, i, s, f"{label} step {step} instance {i}")
    info = ens.info()[:2]
    ens.close()
    for s in sims:
        s.close()
    return info


def test_explosion_ensemble_matches_single_runs():
    """Five explosion instances of different depths (3..6 levels), admesh
    every step: each equals its own single-mesh run bit for bit."""
    cfgs = [_c4(i) for i in C4_PICK]
    n, nleaf = _run_pair(cfgs, 12, "c4")
    assert n == len(cfgs) and nleaf > 0


def test_jet_variant_ensemble_matches_single_runs():
    """Two cooling-jet variants (coolingh, jet source, reflective walls) on a
    reduced 16x4 root layout."""
    cfgs = [P.config_c5(k, nbrootx=16, nbrooty=4, nblocks=4000) for k in range(2)]
    _run_pair(cfgs, 8, "c5")


def test_mixed_outfyb_ensemble_matches_single_runs():
    """An ensemble where only some instances are in the OUTFYB compat mode
    (outflow bottom): every instance's fill I ops go up and are gathered, the
    others' as copies of fill II's; the per-slot uploads follow 12 mesh
    changes of each.  Each instance equals its own single-mesh run."""
    cfgs = [_c4(C4_PICK[0]), _c4(C4_PICK[2]), _c4(C4_PICK[3])]
    cfgs[1].bc = ("OUTFLOW", "OUTFLOW", "REFLECTIVE", "OUTFLOW")
    cfgs[2].bc = ("REFLECTIVE", "OUTFLOW", "REFLECTIVE", "REFLECTIVE")
    _run_pair(cfgs, 12, "mixed outfyb")


def _ens(cfgs, graph=False, groups=1):
    e = Ensemble(cfgs, groups=groups)
    e.basegrid()
    for _ in range(3):
        e.admesh()
    e.set_graph(graph)
    return e


def _same_state(a, b, n, label):
    for i in range(n):
        ta, dta, nba, nla = a.instance(i)
        tb, dtb, nbb, nlb = b.instance(i)
        assert (ta, dta, nba, nla) == (tb, dtb, nbb, nlb), f"{label} instance {i}"
        la = a.tree(i)[1]
        ids = np.nonzero(la)[0]
        assert np.array_equal(a.download(i)[ids], b.download(i)[ids]), f"{label} instance {i}: tiles differ"


def test_graph_replays_match_eager_steps():
    """The captured step (two graphs, one per tile-buffer parity) replayed
    with admesh, table uploads and synchronisations between replays: every
    replay's state equals the eager run's."""
    cfgs = [_c4(i) for i in C4_PICK[:3]]
    a = _ens(cfgs, graph=False)
    b = _ens(cfgs, graph=True)
    for step in range(8):
        a.step(1)
        b.step(1)
        _same_state(a, b, len(cfgs), f"graph step {step}")
    a.close()
    b.close()


def test_autotuned_launch_shapes_keep_results():
    """Tuning after 4 steps, then 6 graph-captured steps with the tuned
    persistent grids: the same states as 10 default eager steps."""
    cfgs = [_c4(i) for i in C4_PICK[:3]]
    a = _ens(cfgs)
    b = _ens(cfgs)
    a.step(4)
    b.step(4)
    grids, threads, us = b.tune(reps=2)
    assert len(grids) == 3 and all(g > 0 for g in grids) and all(t > 0 for t in us)
    assert all(t in (256, 384, 512) for t in threads)
    b.set_graph(True)
    a.step(6)
    b.step(6)
    _same_state(a, b, len(cfgs), "tuned")
    a.close()
    b.close()


@pytest.mark.parametrize("threads", [256, 384, 512])
def test_launch_shapes_keep_results(threads):
    """Stage kernels with far fewer workgroups than leaf blocks (each
    workgroup loops over the queue) and 4, 6 or 8 waves per workgroup give
    the same states as one 4-wave workgroup per block."""
    cfgs = [_c4(i) for i in C4_PICK[:2]]
    a = _ens(cfgs)
    b = _ens(cfgs)
    for k, g in enumerate((7, 13, 5)):
        b.set_launch(k, g, threads)
    a.step(5)
    b.step(5)
    _same_state(a, b, len(cfgs), f"launch shapes {threads}")
    a.close()
    b.close()


def test_jet_launch_shapes_keep_results():
    """The same on the jet (jet source, coolingh, reflective walls) with 6- and
    8-wave workgroups, where stage B gathers its ghost ring in waves of its own."""
    cfgs = [P.config_c5(k, nbrootx=16, nbrooty=4, nblocks=4000) for k in range(2)]
    a = _ens(cfgs)
    b = _ens(cfgs)
    b.set_launch(0, 0, 512)
    b.set_launch(1, 0, 384)
    b.set_launch(2, 0, 384)
    a.step(6)
    b.step(6)
    _same_state(a, b, len(cfgs), "jet launch shapes")
    a.close()
    b.close()


def test_grouped_explosion_ensemble_matches_single_runs():
    """The five explosion instances split into three groups (2, 2, 1) stepped
    concurrently from host threads, each group on its own stream, with graph
    replays: each instance equals its own single-mesh run bit for bit."""
    cfgs = [_c4(i) for i in C4_PICK]
    n, nleaf = _run_pair(cfgs, 10, "c4 grouped", groups=3, graph=True)
    assert n == len(cfgs) and nleaf > 0


def test_grouped_jet_ensemble_matches_ungrouped():
    """Four jet variants in two concurrent groups against the same four in one
    group, over 12 steps with admesh: identical states."""
    cfgs = [P.config_c5(k, nbrootx=16, nbrooty=4, nblocks=4000) for k in range(4)]
    a = _ens(cfgs, graph=True)
    b = _ens(cfgs, graph=True, groups=2)
    for step in range(3):
        a.step(4)
        b.step(4)
        _same_state(a, b, len(cfgs), f"groups step {step}")
    assert a.info()[1] == b.info()[1]
    a.close()
    b.close()


def test_c4_sixty_four_instances():
    """The C4 workload: 64 explosion instances (max levels 3..6, seeded) in
    one ensemble of 4 concurrently stepped groups, 6 steps with admesh;
    instances from every group equal their single-mesh runs bit for bit."""
    cfgs = [P.config_c4(i) for i in range(64)]
    ens = Ensemble(cfgs, groups=4)
    ens.basegrid()
    for _ in range(3):
        ens.admesh()
    ens.set_graph(True)
    picks = [0, 21, 42, 63]
    sims = {}
    for i in picks:
        s = Simulation(cfgs[i])
        s.basegrid()
        for _ in range(3):
            s.admesh()
        sims[i] = s
    for step in range(6):
        ens.step(1)
        for i in picks:
            sims[i].step(1, admesh=cfgs[i].nlevs > 1)
            _compare_instance(ens, i, sims[i], f"c4x64 step {step} instance {i}")
    assert ens.info()[0] == 64
    ens.close()
    for s in sims.values():
        s.close()
