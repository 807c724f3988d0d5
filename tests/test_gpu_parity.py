"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded inputs.

Bar: octree structure, vacancy counts and FP64 centroids bit-exact; target picks identical except
for searches whose draw lies within 1e-12 (relative to the total) of a cumulative boundary in the
oracle; acceptance-test counts per search equal; commit results and vacancies bit-exact; neuron
step x, u, calcium bit-exact and fired flags equal, growth deltas within 1e-9 relative.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from relearn_amd import RL_FLAG_EXACT_ORDER, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _tree_pair(oracle, xyz, signal, box_min, box_max, ve, vi, enabled=None, precision=RL_PREC_FP64):
    t = oracle.Tree(xyz, box_min, box_max)
    t.update(ve, vi, enabled)
    g = BarnesHutHIP(box_min, box_max, precision=precision)
    g.init(xyz, signal)
    if enabled is not None:
        g.set_enabled(enabled)
    g.prepare_update_connectivity(np.zeros(len(xyz), np.uint32), ve, vi)
    return t, g


def _compare_trees(a, b):
    assert len(a["level"]) == len(b["level"])
    for k in ("level", "nchild", "link", "octant", "box_lo", "box_hi", "side", "cnt_exc", "cnt_inh", "has_exc",
              "has_inh", "pos_exc", "pos_inh"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("n,side,seed", [(1, 1000.0, 0), (2, 1000.0, 1), (1000, 1000.0, 0), (200_000, 10_000.0, 2)])
def test_tree_parity_uniform(oracle, n, side, seed):
    _say("tree uniform", n, side, seed)
    net = uniform_cube(n, side, seed)
    t, g = _tree_pair(oracle, net.xyz, net.signal, net.box_min, net.box_max, net.vac_exc, net.vac_inh)
    _compare_trees(t.export(), g.export_tree())
    g.close()


def test_tree_parity_disabled_neurons(oracle):
    _say("tree disabled neurons")
    net = uniform_cube(5000, 2000.0, 3)
    en = (np.random.default_rng(3).random(5000) > 0.25).astype(np.uint8)
    t, g = _tree_pair(oracle, net.xyz, net.signal, net.box_min, net.box_max, net.vac_exc, net.vac_inh, en)
    _compare_trees(t.export(), g.export_tree())
    g.close()


def test_tree_parity_real_positions(oracle):
    _say("tree real positions 51842")
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    n = len(d["xyz"])
    rng = np.random.default_rng(4)
    ve = rng.poisson(1.5, n).astype(np.uint32)
    vi = rng.poisson(1.5, n).astype(np.uint32)
    t, g = _tree_pair(oracle, d["xyz"], d["signal"], d["box_min"], d["box_max"], ve, vi)
    _compare_trees(t.export(), g.export_tree())
    g.close()


def test_tree_rejects_bad_inputs():
    _say("tree rejects bad inputs")
    g = BarnesHutHIP((0, 0, 0), (10, 10, 10))
    with pytest.raises(Exception):
        g.init(np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]]), np.zeros(2, np.uint8))
    with pytest.raises(Exception):
        g.init(np.array([[11.0, 1.0, 1.0]]), np.zeros(1, np.uint8))
    g.close()
    with pytest.raises(Exception):
        BarnesHutHIP((0, 0, 0), (10, 10, 10), theta=0.6)


def _search_pair(oracle, net, seed, step=0, theta=0.3, sigma=750.0, enabled=None, precision=RL_PREC_FP64, flags=0):
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh, enabled)
    ref = t.find_targets(net.signal, net.vac_axon, enabled, theta=theta, sigma=sigma, seed=seed, step=step,
                         threads=0)
    g = BarnesHutHIP(net.box_min, net.box_max, theta=theta, sigma=sigma, seed=seed, precision=precision, flags=flags)
    g.init(net.xyz, net.signal)
    if enabled is not None:
        g.set_enabled(enabled)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    nreq = g.find_target_neurons(step)
    slots = int(net.vac_axon.astype(np.uint64).sum())
    tgt, tests, cnt = g.get_search_slots(slots)
    return ref, g, nreq, tgt, tests, cnt


def _check_search(ref, tgt, tests, cnt, net):
    near = ref["margin"] < 1e-12
    exact_counts = ref["count"] == cnt
    # a neuron whose searches are all clear of boundaries must agree on the number of requests
    offs = ref["offsets"].astype(np.int64)
    n = len(cnt)
    bad = []
    for i in range(n):
        k = int(net.vac_axon[i])
        if k == 0:
            continue
        sl = slice(offs[i], offs[i] + k)
        if near[sl].any():
            continue
        if not exact_counts[i]:
            bad.append(("count", i))
            continue
        c = int(cnt[i])
        if not np.array_equal(ref["target"][sl][:c], tgt[sl][:c]):
            bad.append(("target", i))
        # acceptance tests of every executed search (including a failed last one)
        e = min(c + 1, k)
        if not np.array_equal(ref["node_tests"][sl][:e], tests[sl][:e]):
            bad.append(("tests", i))
    return bad, int(near.sum())


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
@pytest.mark.parametrize("n,side,seed", [(1000, 1000.0, 0), (30_000, 10_000.0 * (0.03 ** (1 / 3)), 5)])
def test_search_parity_fp64(oracle, n, side, seed, flags):
    _say("search fp64", n, side, seed, "flags", flags)
    net = uniform_cube(n, side, seed)
    ref, g, nreq, tgt, tests, cnt = _search_pair(oracle, net, seed=seed + 11, step=3, flags=flags)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("near-boundary searches", near, "mismatches", len(bad), bad[:5])
    assert not bad
    assert nreq == int(cnt.sum())
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_search_parity_fp64_disabled_and_sparse(oracle, flags):
    _say("search fp64 disabled + sparse vacancies", flags)
    net = uniform_cube(4000, 1500.0, 9)
    rng = np.random.default_rng(9)
    en = (rng.random(4000) > 0.2).astype(np.uint8)
    net.vac_exc[rng.random(4000) < 0.7] = 0
    ref, g, nreq, tgt, tests, cnt = _search_pair(oracle, net, seed=21, step=1, enabled=en, flags=flags)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    assert not bad
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_search_parity_real_positions(oracle, flags):
    _say("search fp64 real positions", flags)
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    n = len(d["xyz"])
    rng = np.random.default_rng(6)
    from relearn_amd.workloads import Network

    sig = (rng.random(n) < 0.2).astype(np.uint8)
    net = Network(d["xyz"], sig, rng.poisson(2, n).astype(np.uint32), rng.poisson(1.5, n).astype(np.uint32),
                  rng.poisson(1.5, n).astype(np.uint32), tuple(d["box_min"]), tuple(d["box_max"]))
    ref, g, nreq, tgt, tests, cnt = _search_pair(oracle, net, seed=7, step=0, sigma=20.0, flags=flags)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("near-boundary searches", near, "mismatches", len(bad))
    assert not bad
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
@pytest.mark.parametrize("theta", [0.1, 0.5])
def test_search_parity_other_theta(oracle, theta, flags):
    # the ends of the acceptance parameter's range (BarnesHutBase.h:74-77).  At 0.1 the level-1
    # lists are long: sources whose expanded nodes overflow a wave's slot table go to the per-lane
    # kernel, so that hand-over is checked too.
    _say("search fp64 theta", theta, "flags", flags)
    net = uniform_cube(20_000, 10_000.0 * (0.02 ** (1 / 3)), 29)
    ref, g, nreq, tgt, tests, cnt = _search_pair(oracle, net, seed=31, step=4, theta=theta, flags=flags)
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("near-boundary searches", near, "mismatches", len(bad), bad[:5])
    assert not bad
    assert nreq == int(cnt.sum())
    g.close()


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_search_parity_gaussian_with_mu(oracle, precision):
    # GaussianDistributionKernel with mu != 0 (Gaussian.h:62-67): the attraction peaks at distance mu
    _say("search with mu = 400 um, precision", precision)
    net = uniform_cube(30_000, 10_000.0 * (0.03 ** (1 / 3)), 23)
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    ref = t.find_targets(net.signal, net.vac_axon, mu=400.0, sigma=750.0, seed=17, step=2, threads=0)
    t.close()
    g = BarnesHutHIP(net.box_min, net.box_max, mu=400.0, sigma=750.0, seed=17, precision=precision)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.find_target_neurons(2)
    tgt, tests, cnt = g.get_search_slots(int(net.vac_axon.sum()))
    if precision == RL_PREC_FP64:
        bad, near = _check_search(ref, tgt, tests, cnt, net)
        _say("near-boundary searches", near, "mismatches", len(bad))
        assert not bad
    else:
        valid = (np.asarray(ref["target"]) != 0xFFFFFFFF) & (tgt != 0xFFFFFFFF)
        div = np.mean(np.asarray(ref["target"])[valid] != tgt[valid])
        _say("pick divergence fp32-mixed vs oracle", div)
        assert div <= 1e-3
        # the acceptance decisions are the FP64 ones: where the picks agree, so do the test counts
        same = valid & (np.asarray(ref["target"]) == tgt)
        assert np.array_equal(np.asarray(ref["node_tests"])[same], tests[same])
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_search_fp32_mixed_divergence(oracle, flags):
    _say("search fp32 mixed vs fp64", flags)
    net = uniform_cube(50_000, 10_000.0 * (0.05 ** (1 / 3)), 12)
    outs = {}
    for prec in (RL_PREC_FP64, RL_PREC_FP32_MIXED):
        g = BarnesHutHIP(net.box_min, net.box_max, seed=5, precision=prec, flags=flags)
        g.init(net.xyz, net.signal)
        g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
        g.find_target_neurons(2)
        slots = int(net.vac_axon.sum())
        outs[prec] = g.get_search_slots(slots)
        g.close()
    t64, tests64, c64 = outs[RL_PREC_FP64]
    t32, tests32, c32 = outs[RL_PREC_FP32_MIXED]
    # identical candidate lists: the reference test counts agree
    assert np.array_equal(tests64, tests32)
    valid = t64 != 0xFFFFFFFF
    div = np.mean(t64[valid] != t32[valid])
    _say("pick divergence fp32-mixed vs fp64", div)
    assert div <= 1e-3


def test_commit_parity(oracle):
    _say("commit parity")
    net = uniform_cube(20_000, 3000.0, 14)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=3)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    m = g.find_target_neurons(4)
    tgt, src, sig, jj = g.get_requests(m)
    # requests are ordered by (source, j)
    assert np.all(np.diff(src.astype(np.int64)) >= 0)
    created = g.commit(4)
    ok = g.get_commit_result(m)
    va, ve, vi = g.get_vacancies()
    e, i, a = net.vac_exc.copy(), net.vac_inh.copy(), net.vac_axon.copy()
    ok_ref, created_ref = oracle.commit(tgt, src, sig, jj, e, i, a, seed=3, step=4)
    assert created == created_ref and np.array_equal(ok, ok_ref)
    assert np.array_equal(ve, e) and np.array_equal(vi, i) and np.array_equal(va, a)
    g.close()


def test_commit_parity_explicit_order(oracle):
    _say("commit parity explicit order")
    net = uniform_cube(5000, 1500.0, 15)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=8)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    m = g.find_target_neurons(0)
    tgt, src, sig, jj = g.get_requests(m)
    order = oracle.reference_shuffle(m, 99)
    created = g.commit(0, order=order)
    ok = g.get_commit_result(m)
    va, ve, vi = g.get_vacancies()
    e, i, a = net.vac_exc.copy(), net.vac_inh.copy(), net.vac_axon.copy()
    ok_ref, created_ref = oracle.commit(tgt, src, sig, jj, e, i, a, order=order)
    assert created == created_ref and np.array_equal(ok, ok_ref)
    assert np.array_equal(ve, e) and np.array_equal(vi, i) and np.array_equal(va, a)
    g.close()


def _izh(oracle_mod):
    p = oracle_mod.StepParams()
    p.a, p.b, p.c, p.d = 0.02, 0.2, -65.0, 8.0
    p.v_spike, p.k1, p.k2, p.k3 = 30.0, 0.04, 5.0, 140.0
    p.h = 10
    p.tau_ca, p.beta_ca = 10000.0, 0.001
    p.eta[0], p.eta[1], p.eta[2] = 0.4, 0.1, 0.0
    p.nu[0], p.nu[1], p.nu[2] = 1e-4, 1e-4, 1e-4
    return p


@pytest.mark.parametrize("targets", ["uniform", "per-neuron"])
def test_neuron_step_parity(oracle, targets):
    # one target calcium for every neuron (the step then reads no per-neuron target) or one each
    _say("neuron step parity", targets)
    from relearn_amd import StepParams

    n, steps = 64, 10
    rng = np.random.default_rng(16)
    po = _izh(oracle)
    pg = StepParams()
    for f in ("a", "b", "c", "d", "v_spike", "k1", "k2", "k3", "h", "tau_ca", "beta_ca"):
        setattr(pg, f, getattr(po, f))
    for k in range(3):
        pg.eta[k] = po.eta[k]
        pg.nu[k] = po.nu[k]
    pg.input_mode = 0
    x = rng.uniform(-70, 20, n)
    u = rng.uniform(-15, 5, n)
    ca = rng.uniform(0, 1, n)
    tca = np.full(n, 0.7) if targets == "uniform" else rng.uniform(0.05, 0.9, n)
    en = np.ones(n, np.uint8)
    en[::9] = 0
    g = BarnesHutHIP((0, 0, 0), (1, 1, 1))
    g.init(rng.random((n, 3)), np.zeros(n, np.uint8))
    g.set_enabled(en)
    g.neurons_init(pg, tca, x=x, u=u, ca=ca)
    xo, uo, co = x.copy(), u.copy(), ca.copy()
    fo = np.zeros(n, np.uint8)
    do = np.zeros(3 * n)
    for s in range(steps):
        inp = rng.normal(5.0, 2.0, n)
        oracle.neuron_step(po, xo, uo, co, inp, fo, tca, do, en)
        g.neurons_step(s, inp)
    st = g.neurons_get()
    assert np.array_equal(st["x"], xo) and np.array_equal(st["u"], uo) and np.array_equal(st["ca"], co)
    assert np.array_equal(st["fired"][en == 1], fo[en == 1])
    assert np.allclose(st["delta"].ravel(), do, rtol=1e-9, atol=1e-300)
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_full_plasticity_step_invariants(flags):
    _say("full plasticity step invariants 200k", flags)
    net = uniform_cube(200_000, 10_000.0 * (0.2 ** (1 / 3)), 17)
    for prec in (RL_PREC_FP64, RL_PREC_FP32_MIXED):
        g = BarnesHutHIP(net.box_min, net.box_max, seed=1, precision=prec, flags=flags)
        g.init(net.xyz, net.signal)
        g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
        st = g.update_connectivity(1)
        va, ve, vi = g.get_vacancies()
        # every created synapse consumed one vacant axon and one vacant dendrite
        assert int(net.vac_axon.sum()) - int(va.sum()) == st["created"]
        assert int(net.vac_exc.sum() + net.vac_inh.sum()) - int(ve.sum() + vi.sum()) == st["created"]
        assert np.all(va <= net.vac_axon) and np.all(ve <= net.vac_exc) and np.all(vi <= net.vac_inh)
        assert st["requests"] <= st["searches"] <= int(net.vac_axon.sum())
        g.close()


def test_coop_matches_exact_order_kernel():
    _say("cooperative vs exact-order kernel 100k")
    net = uniform_cube(100_000, 10_000.0 * (0.1 ** (1 / 3)), 18)
    outs = {}
    for flags in (0, RL_FLAG_EXACT_ORDER):
        g = BarnesHutHIP(net.box_min, net.box_max, seed=9, flags=flags)
        g.init(net.xyz, net.signal)
        g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
        g.find_target_neurons(5)
        outs[flags] = g.get_search_slots(int(net.vac_axon.sum())) + (g.last_search_stats(),)
        g.close()
    a, b = outs[0], outs[RL_FLAG_EXACT_ORDER]
    assert np.array_equal(a[1], b[1])          # reference test counts per search
    div = np.mean(a[0] != b[0])
    _say("cooperative vs exact-order pick difference", div)
    assert div < 1e-4
    assert a[3]["node_tests_ref"] == b[3]["node_tests_ref"]


def _check_search_vec(ref, tgt, tests, cnt, net):
    """Vectorized form of _check_search for large networks."""
    va = net.vac_axon.astype(np.int64)
    slots = int(va.sum())
    sn = np.repeat(np.arange(len(va)), va)
    j = np.arange(slots) - ref["offsets"].astype(np.int64)[sn]
    near_n = np.zeros(len(va), bool)
    np.logical_or.at(near_n, sn, ref["margin"] < 1e-12)
    ok_n = ~near_n & (va > 0)
    cnt_bad = ok_n & (ref["count"] != cnt)
    c = cnt.astype(np.int64)[sn]
    chk = ok_n[sn] & ~cnt_bad[sn]
    t_bad = chk & (j < c) & (ref["target"] != tgt)
    e_bad = chk & (j < np.minimum(c + 1, va[sn])) & (ref["node_tests"] != tests)
    return int(cnt_bad.sum()), int(t_bad.sum()), int(e_bad.sum()), int(near_n.sum())


def test_million_neurons_fp64(oracle):
    """1M neurons uniform in a 10,000 um cube, Poisson(2)/(1.5)/(1.5) vacancies, theta 0.3, sigma 750:
    bit-exact octree counts and centroids, identical picks except draws within 1e-12 of a boundary."""
    _say("1M neurons fp64 tree + search parity")
    net = uniform_cube(1_000_000, 10_000.0, 0)
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=1, precision=RL_PREC_FP64)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    a, b = t.export(), g.export_tree()
    _compare_trees(a, b)
    _say("1M-neuron tree identical", len(a["level"]), "nodes")
    ref = t.find_targets(net.signal, net.vac_axon, theta=0.3, sigma=750.0, seed=1, step=0, threads=0)
    nreq = g.find_target_neurons(0)
    slots = int(net.vac_axon.astype(np.uint64).sum())
    tgt, tests, cnt = g.get_search_slots(slots)
    cb, tb, eb, near = _check_search_vec(ref, tgt, tests, cnt, net)
    _say("1M near-boundary neurons", near, "count/target/test mismatches", cb, tb, eb)
    assert cb == 0 and tb == 0 and eb == 0
    assert near <= 1e-3 * len(cnt)
    assert nreq == int(cnt.sum())
    g.close()
    t.close()
