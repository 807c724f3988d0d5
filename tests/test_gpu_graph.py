"""Captured plasticity periods (rl_bh_graph_create / rl_bh_graph_launch) against the eager entry points.

A replay runs the same kernels as the eager calls, with the update index and the neuron steps read
from device memory, so every replay must reproduce the eager results bit for bit: requests, accept
flags, vacancies, counters and the neuron state.  The eager path is itself checked against the CPU
oracle in test_gpu_parity.py.  Each graph is replayed several times with eager work (uploads, reads,
synchronisations) between the replays, and every replay is compared.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_FLAG_EXACT_ORDER, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu

RL_ELIMIT = -7


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _update_result(ctx, stats):
    m = stats["requests"]
    t, s, g, j = ctx.get_requests(m)
    ok = ctx.get_commit_result(m)
    return dict(t=t, s=s, g=g, j=j, ok=ok, vac=ctx.get_vacancies(), stats=stats)


def _same(a, b):
    for k in ("t", "s", "g", "j", "ok"):
        assert np.array_equal(a[k], b[k]), k
    for t in range(3):
        assert np.array_equal(a["vac"][t], b["vac"][t]), ("vac", t)
    for k in ("searches", "requests", "created", "node_tests", "node_tests_ref"):
        assert a["stats"][k] == b["stats"][k], k


@pytest.mark.parametrize("precision,flags", [(RL_PREC_FP64, 0), (RL_PREC_FP32_MIXED, 0),
                                             (RL_PREC_FP64, RL_FLAG_EXACT_ORDER),
                                             (RL_PREC_FP32_MIXED, RL_FLAG_EXACT_ORDER)])
def test_graph_update_matches_eager(precision, flags):
    _say("graph connectivity update vs eager, 20000 neurons, precision", precision, "flags", flags)
    net = uniform_cube(20_000, 3000.0, 31)
    rng = np.random.default_rng(5)
    eager = BarnesHutHIP(net.box_min, net.box_max, seed=11, precision=precision, flags=flags)
    graph = BarnesHutHIP(net.box_min, net.box_max, seed=11, precision=precision, flags=flags)
    for c in (eager, graph):
        c.init(net.xyz, net.signal)
    cap = int(2 * net.vac_axon.sum())
    graph.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    graph.graph_create(0, cap)
    for rep, step in enumerate((3, 3, 7, 0)):
        va = rng.poisson(2.0, net.n).astype(np.uint32)
        ve = rng.poisson(1.5, net.n).astype(np.uint32)
        vi = rng.poisson(1.5, net.n).astype(np.uint32)
        eager.prepare_update_connectivity(va, ve, vi)
        E = _update_result(eager, eager.update_connectivity(step))
        graph.prepare_update_connectivity(va, ve, vi)
        G = _update_result(graph, graph.graph_launch(0, step))
        _say("replay", rep, "update", step, "requests", G["stats"]["requests"], "created", G["stats"]["created"])
        assert G["stats"]["created"] > 0
        _same(E, G)
    eager.close()
    graph.close()


def _neuron_setup(ctx, n, seed, nu):
    rng = np.random.default_rng(seed)
    p = StepParams.izhikevich(input_mode=1, nu=nu)
    grown = rng.poisson(2.0, (3, n)).astype(np.float64)
    ca = rng.uniform(0.3, 0.6, n)
    ctx.neurons_init(p, np.full(n, 0.7), ca=ca, grown=grown.ravel())


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_graph_period_matches_eager(precision):
    n, steps, periods = 12_000, 25, 3
    _say("graph plasticity period vs eager,", n, "neurons,", steps, "steps x", periods, "periods, precision", precision)
    net = uniform_cube(n, 2500.0, 32)
    eager = BarnesHutHIP(net.box_min, net.box_max, seed=21, precision=precision)
    graph = BarnesHutHIP(net.box_min, net.box_max, seed=21, precision=precision)
    for c in (eager, graph):
        c.init(net.xyz, net.signal)
        _neuron_setup(c, n, 9, 0.02)
    graph.graph_create(steps, 8 * n)
    total_created = 0
    for per in range(periods):
        first = 1000 + per * steps
        for k in range(steps):
            eager.neurons_step(first + k)
        eager.neurons_commit_elements()
        E = _update_result(eager, eager.update_connectivity(per))
        G = _update_result(graph, graph.graph_launch(first, per))
        _say("period", per, "requests", G["stats"]["requests"], "created", G["stats"]["created"])
        _same(E, G)
        total_created += G["stats"]["created"]
        se, sg = eager.neurons_get(), graph.neurons_get()
        for key in ("x", "u", "ca", "fired", "delta", "grown", "connected"):
            assert np.array_equal(se[key], sg[key]), (per, key)
    assert total_created > 0
    eager.close()
    graph.close()


def test_graph_over_capacity_changes_nothing():
    _say("graph over capacity")
    net = uniform_cube(5000, 1500.0, 33)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=2)
    g.init(net.xyz, net.signal)
    small = int(net.vac_axon.sum()) // 2
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.graph_create(0, small)
    with pytest.raises(BHError) as e:
        g.graph_launch(0, 1)
    assert e.value.code == RL_ELIMIT
    va, ve, vi = g.get_vacancies()
    assert np.array_equal(va, net.vac_axon) and np.array_equal(ve, net.vac_exc) and np.array_equal(vi, net.vac_inh)
    # an asynchronous replay that overflows is reported by the next replay with stats, then cleared
    g.graph_launch(0, 1, stats=False)
    half = (net.vac_axon // 4).astype(np.uint32)
    g.prepare_update_connectivity(half, net.vac_exc, net.vac_inh)
    with pytest.raises(BHError):
        g.graph_launch(0, 1)
    g.prepare_update_connectivity(half, net.vac_exc, net.vac_inh)
    st = g.graph_launch(0, 1)
    assert st["created"] > 0
    g.close()


def test_graph_ensemble_period_matches_eager():
    from relearn_amd.workloads import parameter_sweep

    ens = parameter_sweep(num_networks=12, n_min=500, n_max=5000, seed=4)
    n, steps = ens.n, 15
    _say("graph ensemble period vs eager,", len(ens.offsets) - 1, "networks,", n, "neurons")
    lo, hi = ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0)
    eager = BarnesHutHIP(lo, hi, seed=5, precision=RL_PREC_FP32_MIXED)
    graph = BarnesHutHIP(lo, hi, seed=5, precision=RL_PREC_FP32_MIXED)
    nu = np.repeat(ens.nu * 200.0, 3)
    for c in (eager, graph):
        c.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
        _neuron_setup(c, n, 12, 0.01)
        c.neurons_set_network_growth(nu)
    graph.graph_create(steps, 8 * n)
    for per in range(2):
        for k in range(steps):
            eager.neurons_step(per * steps + k)
        eager.neurons_commit_elements()
        E = _update_result(eager, eager.update_connectivity(per))
        G = _update_result(graph, graph.graph_launch(per * steps, per))
        _say("period", per, "created", G["stats"]["created"])
        _same(E, G)
        se, sg = eager.neurons_get(), graph.neurons_get()
        for key in ("x", "u", "ca", "grown", "connected"):
            assert np.array_equal(se[key], sg[key]), (per, key)
    eager.close()
    graph.close()


def _pairs(off, src, w):
    t = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    out = {}
    for a, b, c in zip(t.tolist(), src.tolist(), w.tolist()):
        out[(a, b)] = out.get((a, b), 0) + c
    return {k: v for k, v in out.items() if v != 0}


def test_graph_ensemble_with_synaptic_input_and_deletion_matches_eager():
    # an ensemble whose periods push synaptic input over the graph the updates build, delete the
    # synapses of shrinking elements and outgrow the graph's initial capacity
    from relearn_amd.workloads import parameter_sweep

    ens = parameter_sweep(num_networks=12, n_min=500, n_max=5000, seed=6)
    n, steps, periods = ens.n, 15, 4
    _say("graph ensemble period with synaptic input and deletion vs eager,", n, "neurons")
    lo, hi = ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0)
    ctxs = []
    for _ in range(2):
        c = BarnesHutHIP(lo, hi, seed=5, precision=RL_PREC_FP32_MIXED)
        c.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
        c.net_init(n // 4)
        rng = np.random.default_rng(13)
        p = StepParams.izhikevich(input_mode=1, nu=0.05)
        p.synaptic_input, p.synapse_conductance = 1, 2.0
        for t in range(3):
            p.eta[t] = (0.02, 0.01, 0.0)[t]
        # targets on both sides of the calcium range: some neurons grow elements, others shrink
        c.neurons_init(p, rng.uniform(0.05, 0.15, n), ca=rng.uniform(0.05, 0.15, n),
                       grown=rng.poisson(2.0, (3, n)).astype(np.float64).ravel())
        c.neurons_set_network_growth(np.repeat(ens.nu * 200.0, 3))
        ctxs.append(c)
    eager, graph = ctxs
    graph.graph_create(steps, 8 * n)
    deleted = regrown = 0
    for per in range(periods):
        for k in range(steps):
            eager.neurons_step(per * steps + k)
        eager.neurons_commit_elements(per)
        E = _update_result(eager, eager.update_connectivity(per))
        G = _update_result(graph, graph.graph_launch(per * steps, per))
        deleted += G["stats"]["deleted"]
        regrown += G["stats"]["net_regrown"]
        _say("period", per, "created", G["stats"]["created"], "deleted", G["stats"]["deleted"],
             "net regrown", G["stats"]["net_regrown"])
        _same(E, G)
        se, sg = eager.neurons_get(), graph.neurons_get()
        for key in ("x", "u", "ca", "fired", "delta", "grown", "connected"):
            assert np.array_equal(se[key], sg[key]), (per, key)
        assert _pairs(*eager.net_get_edges()) == _pairs(*graph.net_get_edges()), per
    assert regrown > 0
    _say("deleted over the run", deleted)
    eager.close()
    graph.close()


def test_graph_needs_counter_input():
    net = uniform_cube(100, 100.0, 34)
    g = BarnesHutHIP(net.box_min, net.box_max)
    g.init(net.xyz, net.signal)
    g.neurons_init(StepParams.izhikevich(input_mode=0), np.full(100, 0.7))
    with pytest.raises(BHError):
        g.graph_create(10, 1000)
    with pytest.raises(BHError):
        g.graph_launch(0, 0)
    g.graph_create(0, 1000)  # the connectivity update alone needs no neuron input
    g.close()


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_search_tuning_does_not_change_results(precision):
    _say("search tuning invariance, precision", precision)
    net = uniform_cube(30_000, 3000.0, 35)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=6, precision=precision)
    g.init(net.xyz, net.signal)
    slots = int(net.vac_axon.sum())
    ref = None
    for w, k in ((1, 1), (5, 1), (16, 1), (24, 1), (16, 4), (3, 64)):
        g.set_tuning(w, k)
        tu = g.get_tuning()
        assert (tu["search_waves_per_cu"], tu["search_chunk"]) == (w, k)
        g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
        g.find_target_neurons(2)
        cur = g.get_search_slots(slots)
        if ref is None:
            ref = cur
        for a, b in zip(ref, cur):
            assert np.array_equal(a, b), (w, k)
    best = g.autotune(2)
    if precision == RL_PREC_FP32_MIXED:
        # 30,000 neurons: a wavefront per source is faster than the group walk, so the walk's setting stays
        # as set and the walk's threshold stays above this update
        assert (best["search_waves_per_cu"], best["search_chunk"]) == (3, 64)
        assert best["walk_min_sources"] > 30_000
    else:
        assert 1 <= best["search_waves_per_cu"] <= 24 and 1 <= best["search_chunk"] <= 8
    for a, b in zip(ref, g.get_search_slots(slots)):
        assert np.array_equal(a, b)
    with pytest.raises(BHError):
        g.set_tuning(0)
    with pytest.raises(BHError):
        g.set_tuning(16, 65)
    g.close()


def test_graph_without_stats_refuses_when_the_network_graph_may_overflow():
    # a replay without stats cannot grow the network graph, so it runs only while the graph holds the
    # most synapses the period can create; a replay with stats grows it and completes the period
    _say("graph without stats, network graph capacity")
    net = uniform_cube(3000, 1400.0, 41)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=2)
    g.init(net.xyz, net.signal)
    g.net_init(16)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.graph_create(0, int(net.vac_axon.sum()))
    with pytest.raises(BHError) as e:
        g.graph_launch(0, 1, stats=False)
    assert e.value.code == RL_ELIMIT
    va, _, _ = g.get_vacancies()
    assert np.array_equal(va, net.vac_axon)   # nothing ran
    st = g.graph_launch(0, 1)
    assert st["created"] > 0 and st["net_regrown"] == 1
    off, src, w = g.net_get_edges()[:3]
    assert int(np.abs(w).sum()) == st["created"]
    # with room for a whole period, a replay without stats runs
    g.net_init(int(net.vac_axon.sum()) + 64)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.graph_create(0, int(net.vac_axon.sum()))
    g.graph_launch(0, 2, stats=False)
    st = g.graph_launch(0, 3)
    g.close()


@pytest.mark.parametrize("n", [30_000, 4_000_000])
def test_autotune_moves_the_walk_threshold_to_the_faster_kernel(n):
    # FP32-mixed: the autotuner times the group walk and a wavefront per source on the prepared update
    # and moves the threshold so that an update of this size takes the faster one: at 30,000 neurons the
    # wavefront per source (the threshold rises above the update), at 4M the walk (it falls to the
    # update's active sources)
    _say("autotuned traversal width,", n, "neurons")
    net = uniform_cube(n, 10_000.0 * (n / 1e6) ** (1 / 3), 47)
    active = int((net.vac_axon > 0).sum())
    g = BarnesHutHIP(net.box_min, net.box_max, seed=6, precision=RL_PREC_FP32_MIXED)
    g.init(net.xyz, net.signal)
    g.set_tuning(24, 1, 1 if n < 1_000_000 else 10 ** 12)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    best = g.autotune(1)
    _say("  tuning", best, "active sources", active)
    assert best["walk_min_sources"] == (active + 1 if n < 1_000_000 else active)
    assert g.get_tuning() == best
    g.close()
