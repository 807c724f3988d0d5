"""The FP32-mixed group walk (one lane per source, 64 neighbouring sources per wavefront) against the
CPU oracle.

With sigma = 10^12 um every Gaussian factor is exactly 1 in FP64 (exp(-d^2/sigma^2) = exp(-1e-18)) and
in the FP32 kernel (exp2 of a number of magnitude 1e-18), so every candidate's probability is its
vacant count in both precisions and every running sum is an exact integer.  The FP32-mixed picks must
then be the oracle's FP64 picks exactly, with the same acceptance-test counts: this pins the walk's
traversal, its table, the top-down resolution of the draws and the descents, for group sizes from 1
to many groups, disabled neurons, empty lists, clustered positions and long lists (theta 0.1, where
groups overflow the table and go to the per-lane kernel).  These tests force the walk
(RL_FLAG_WALK_ALWAYS); by default the device takes it only for updates with many active sources.  With the Gaussian at sigma = 750 um the
FP32 probabilities differ from FP64 in the last bits: the picks may differ only in a small fraction of
searches, the test counts must agree wherever the picks do, and the walk must agree with the
wave-per-source FP32 search (RL_FLAG_WAVE_PER_SOURCE) just as closely.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from relearn_amd import (RL_FLAG_COUNT_WALK, RL_FLAG_WALK_ALWAYS, RL_FLAG_WAVE_PER_SOURCE, RL_PREC_FP32_MIXED,
                         RL_PREC_FP64, BarnesHutHIP)
from relearn_amd.workloads import Network, uniform_cube
from tests.helpers import deep_cluster
from tests.test_gpu_parity import _check_search

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FLAT = 1.0e12


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _pair(oracle, net, seed, step, sigma, theta=0.3, enabled=None, flags=RL_FLAG_WALK_ALWAYS, precision=RL_PREC_FP32_MIXED):
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh, enabled)
    ref = t.find_targets(net.signal, net.vac_axon, enabled, theta=theta, sigma=sigma, seed=seed, step=step, threads=0)
    t.close()
    g = BarnesHutHIP(net.box_min, net.box_max, theta=theta, sigma=sigma, seed=seed, precision=precision, flags=flags)
    g.init(net.xyz, net.signal)
    if enabled is not None:
        g.set_enabled(enabled)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    nreq = g.find_target_neurons(step)
    st = g.last_search_stats()
    tgt, tests, cnt = g.get_search_slots(int(net.vac_axon.astype(np.uint64).sum()))
    g.close()
    return ref, nreq, st, tgt, tests, cnt


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 130, 1000, 20_000])
def test_walk_flat_kernel_matches_oracle_exactly(oracle, n):
    side = 1000.0 * max(n / 1000.0, 0.01) ** (1 / 3)
    net = uniform_cube(n, side, 40 + n % 7)
    _say("walk, flat kernel,", n, "neurons")
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=3, step=5, sigma=FLAT)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("  near-boundary", near, "mismatches", len(bad), bad[:5], "lane sources", st["lane_sources"])
    assert not bad
    assert nreq == int(cnt.sum())


@pytest.mark.parametrize("theta", [0.1, 0.5])
def test_walk_flat_kernel_other_theta(oracle, theta):
    # at 0.1 the lists are long enough that groups overflow the walk's table and go to the per-lane kernel
    net = uniform_cube(20_000, 10_000.0 * (0.02 ** (1 / 3)), 29)
    _say("walk, flat kernel, theta", theta)
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=31, step=4, sigma=FLAT, theta=theta)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("  near-boundary", near, "mismatches", len(bad), bad[:5], "lane sources", st["lane_sources"])
    assert not bad


def test_walk_flat_kernel_disabled_and_sparse(oracle):
    net = uniform_cube(6000, 1600.0, 9)
    rng = np.random.default_rng(9)
    en = (rng.random(6000) > 0.2).astype(np.uint8)
    net.vac_exc[rng.random(6000) < 0.7] = 0
    net.vac_inh[rng.random(6000) < 0.95] = 0
    _say("walk, flat kernel, disabled neurons and sparse vacancies")
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=21, step=1, sigma=FLAT, enabled=en)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    assert not bad


def test_walk_flat_kernel_real_positions(oracle):
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    n = len(d["xyz"])
    rng = np.random.default_rng(6)
    sig = (rng.random(n) < 0.2).astype(np.uint8)
    net = Network(d["xyz"], sig, rng.poisson(2, n).astype(np.uint32), rng.poisson(1.5, n).astype(np.uint32),
                  rng.poisson(1.5, n).astype(np.uint32), tuple(d["box_min"]), tuple(d["box_max"]))
    _say("walk, flat kernel, real positions")
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=7, step=0, sigma=FLAT)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("  near-boundary", near, "mismatches", len(bad), bad[:5])
    assert not bad


def test_walk_gaussian_against_fp64_and_wave_per_source(oracle):
    net = uniform_cube(60_000, 10_000.0 * (0.06 ** (1 / 3)), 12)
    _say("walk, Gaussian sigma 750 um, 60000 neurons")
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=5, step=2, sigma=750.0)
    valid = (ref["target"] != 0xFFFFFFFF) & (tgt != 0xFFFFFFFF)
    div = float(np.mean(ref["target"][valid] != tgt[valid]))
    same = ref["target"] == tgt
    _say("  divergence vs FP64 oracle", div)
    assert div <= 1e-4
    assert np.array_equal(ref["node_tests"][same & valid], tests[same & valid])
    _, _, _, tw, testw, cntw = _pair(oracle, net, seed=5, step=2, sigma=750.0, flags=RL_FLAG_WAVE_PER_SOURCE)
    dw = float(np.mean(tw != tgt))
    _say("  divergence vs the wave-per-source FP32 search", dw)
    assert dw <= 1e-4
    both = tw == tgt
    assert np.array_equal(testw[both], tests[both])


def test_walk_counting_instantiation_same_picks(oracle):
    # RL_FLAG_COUNT_WALK runs a separate instantiation of the walk that also counts its node steps and
    # table slots: the picks and test counts are the plain walk's, each step tests the node for at
    # least one lane and for at most 64, and every expanded node takes one slot
    net = uniform_cube(30_000, 10_000.0 * (0.03 ** (1 / 3)), 17)
    _say("walk, counting instantiation, 30000 neurons")
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=8, step=3, sigma=750.0)
    _, nreq_c, st_c, tgt_c, tests_c, cnt_c = _pair(oracle, net, seed=8, step=3, sigma=750.0,
                                                   flags=RL_FLAG_COUNT_WALK | RL_FLAG_WALK_ALWAYS)
    assert st["walk_steps"] == 0 and st["walk_slots"] == 0 and st["walk_table_bytes"] == 0
    assert np.array_equal(tgt, tgt_c) and np.array_equal(tests, tests_c) and np.array_equal(cnt, cnt_c)
    assert nreq == nreq_c
    _say("  steps", st_c["walk_steps"], "slots", st_c["walk_slots"], "lane sources", st_c["lane_sources"])
    assert st_c["lane_sources"] == 0
    # a lane's level-1 tests are the walk steps its mask took part in (1 to 64 lanes per step); the
    # descents below the picked level-1 node add at most 8 tests per level
    assert st_c["walk_steps"] > 0 and st_c["walk_slots"] > 0
    assert st_c["walk_slots"] <= st_c["walk_steps"]
    # stored: at most the 64 running sums and one 32-byte child-entry row per slot, at least the root's row
    assert 32 <= st_c["walk_table_bytes"] <= (64 * 8 + 32) * st_c["walk_slots"]
    assert st_c["walk_table_bytes"] % 8 == 0
    assert st_c["walk_steps"] <= st_c["node_tests"] <= 64 * st_c["walk_steps"] + 8 * 32 * st_c["searches"]


@pytest.mark.parametrize("flags", [RL_FLAG_WALK_ALWAYS, RL_FLAG_WAVE_PER_SOURCE])
def test_fp32_guard_tiny_cells_flat_kernel(oracle, flags):
    # 1,500 neurons in a 0.05 um cube inside a 25,000 um box: the cluster's cells are a few hundred
    # fixed-point units wide, where rounding the source and the centroid to the grid moves
    # theta^2 d^2 by far more than the relative guard band of the FP32 test.  The band's absolute part
    # sends those decisions to FP64, so the candidate lists, and with the flat kernel the picks and
    # test counts, are the FP64 oracle's in both FP32-mixed kernels.
    rng = np.random.default_rng(11)
    xyz = np.concatenate([rng.random((500, 3)) * 25_000.0, 12_000.0 + rng.random((1500, 3)) * 0.05])
    n = len(xyz)
    net = Network(xyz, (rng.random(n) < 0.2).astype(np.uint8), rng.poisson(2, n).astype(np.uint32),
                  rng.poisson(1.5, n).astype(np.uint32), rng.poisson(1.5, n).astype(np.uint32),
                  (0.0, 0.0, 0.0), (25_000.0, 25_000.0, 25_000.0))
    _say("FP32 guard band, 0.05 um cluster, flags", flags)
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=13, step=2, sigma=FLAT, flags=flags)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("  near-boundary", near, "mismatches", len(bad), bad[:5], "lane sources", st["lane_sources"])
    assert not bad
    assert nreq == int(cnt.sum())


def _gpu_only(net, seed, step, flags):
    g = BarnesHutHIP(net.box_min, net.box_max, theta=0.3, sigma=750.0, seed=seed, precision=RL_PREC_FP32_MIXED,
                     flags=flags)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.find_target_neurons(step)
    st = g.last_search_stats()
    out = g.get_search_slots(int(net.vac_axon.astype(np.uint64).sum()))
    g.close()
    return st, out


@pytest.mark.parametrize("n", [20_000, 4_000_000])
def test_default_kernel_choice_by_active_sources(n):
    # by default the device takes the group walk from 192 x 64 active sources per CU on (3,145,728 on
    # 256 CUs) and a wavefront per source below; the counting flag shows whether the walk ran, and the
    # results are those of the forced choice, bit for bit
    net = uniform_cube(n, 10_000.0 * (n / 1e6) ** (1 / 3), 23)
    _say("default kernel choice,", n, "neurons")
    st, (tgt, tests, cnt) = _gpu_only(net, 4, 1, RL_FLAG_COUNT_WALK)
    walked = st["walk_steps"] > 0
    assert walked == (n > 1_000_000)   # 4M neurons: 3.46M active sources
    forced = RL_FLAG_WALK_ALWAYS if walked else RL_FLAG_WAVE_PER_SOURCE
    st_f, (tgt_f, tests_f, cnt_f) = _gpu_only(net, 4, 1, forced)
    assert np.array_equal(tgt, tgt_f) and np.array_equal(tests, tests_f) and np.array_equal(cnt, cnt_f)


@pytest.mark.parametrize("flags,precision", [(RL_FLAG_WALK_ALWAYS, RL_PREC_FP32_MIXED),
                                             (RL_FLAG_WAVE_PER_SOURCE, RL_PREC_FP32_MIXED),
                                             (0, RL_PREC_FP64)])
def test_deep_cluster_flat_kernel(oracle, flags, precision):
    # a union walk deeper than the LDS ring of staged children (the lists reach level 27 or deeper:
    # test_deep_cluster_lists_reach_past_the_staging_ring): when the walk returns to a level whose block
    # a deeper level took, it copies that level's children again, with the level's own child count
    net = deep_cluster(19)
    _say("deep cluster, flat kernel, flags", flags, "precision", precision)
    ref, nreq, st, tgt, tests, cnt = _pair(oracle, net, seed=17, step=3, sigma=FLAT, flags=flags, precision=precision)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("  near-boundary", near, "mismatches", len(bad), bad[:5], "lane sources", st["lane_sources"])
    assert not bad
    assert nreq == int(cnt.sum())
