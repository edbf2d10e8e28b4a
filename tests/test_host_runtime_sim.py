"""The product's host runtime (walicxe2d_amd/csrc/wx_host.cpp: admesh tree
logic, table compilation, staging, the device table builder's uploads and
launches) built for the CPU against a host-only HIP stand-in with
ASan/UBSan, driven with seeded refinement flags and compared with the
oracle's admesh after every round; the tables it holds (in each builder
mode in turn) against a host compilation of its tree (tests/hostsim)."""
import os
import shutil
import subprocess

import pytest

from _hostsim_build import make as make_hostsim

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostsim")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
@pytest.mark.parametrize("seed", [1, 7])
def test_host_tree_matches_oracle_under_asan(seed):
    make_hostsim()
    r = subprocess.run([os.path.join(HERE, "hostsim_main"), "250", str(seed)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "trees identical, tables identical" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
@pytest.mark.parametrize("groups", [1, 3])
def test_ensemble_trees_match_oracle_under_asan(groups):
    """Three meshes of different root layouts, depths and capacities advanced as
    one ensemble (worker threads on), in one group or in three groups stepped
    from concurrent host threads, with the steps replayed from captured graphs
    (re-captured as the queues grow and shrink, and after launch-shape
    changes): every instance's tree after every admesh equals the oracle's for
    the same flags, no allocation a live graph references is freed, and no
    copy or memset is captured."""
    make_hostsim()
    r = subprocess.run([os.path.join(HERE, "hostsim_main"), "120", "5", "ens", str(groups)], capture_output=True,
                       text=True, timeout=600, env={**os.environ, "WX_HOST_THREADS": "4"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "trees identical" in r.stdout
    assert " 0 graph captures" not in r.stdout and " 0 replays" not in r.stdout, r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
@pytest.mark.parametrize("groups", [2, 3])
def test_grouped_ensemble_under_tsan(groups):
    """The grouped ensemble under ThreadSanitizer: the groups' host threads
    and their worker pools share no unsynchronised state."""
    make_hostsim()
    r = subprocess.run([os.path.join(HERE, "hostsim_tsan"), "40", "3", "ens", str(groups)], capture_output=True,
                       text=True, timeout=600, env={**os.environ, "WX_HOST_THREADS": "6"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "trees identical" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_ghost_tables_match_sequential_list_semantics():
    """The compiled gather tables of both fills (fill II compiled once, fill I
    derived from it plus the OUTFYB writes) equal a direct application of the
    ordered list, entry by entry, on 320 random meshes covering the jet
    source, the OUTFYB compat mode and both boundary kinds (ASan build)."""
    make_hostsim("tables_check")
    r = subprocess.run([os.path.join(HERE, "tables_check")], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "tables identical" in r.stdout
