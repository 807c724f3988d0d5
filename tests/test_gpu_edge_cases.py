"""Degenerate inputs of the connectivity update, each against the oracle or a closed form:
no neurons, no vacancies, no vacant dendrites (every search fails at once), every neuron disabled, a
single neuron (the root is its own leaf), two neurons, neurons on the box's faces and corners, and
one neuron with thousands of vacant axons (many rounds of 8 draws, long candidate lists shared by
all of its axons).  Both search kernels (cooperative and exact-order) and both precisions where the
case allows.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_FLAG_EXACT_ORDER, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
from relearn_amd.workloads import Network, uniform_cube

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _update(net, flags=0, precision=RL_PREC_FP64, enabled=None, seed=3, step=1):
    g = BarnesHutHIP(net.box_min, net.box_max, seed=seed, precision=precision, flags=flags)
    g.init(net.xyz, net.signal)
    if enabled is not None:
        g.set_enabled(enabled)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    st = g.update_connectivity(step)
    return g, st


def _net(xyz, va, ve, vi, sig=None, box=((0, 0, 0), (1000, 1000, 1000))):
    n = len(xyz)
    sig = np.zeros(n, np.uint8) if sig is None else np.asarray(sig, np.uint8)
    return Network(np.asarray(xyz, np.float64).reshape(n, 3), sig, np.asarray(va, np.uint32), np.asarray(ve, np.uint32),
                   np.asarray(vi, np.uint32), box[0], box[1])


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_no_neurons(flags):
    _say("no neurons", flags)
    net = _net(np.zeros((0, 3)), [], [], [])
    g, st = _update(net, flags)
    assert st["searches"] == 0 and st["requests"] == 0 and st["created"] == 0
    g.close()


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_no_vacancies_and_all_disabled(precision, flags):
    _say("no vacancies / all disabled", precision, flags)
    net = uniform_cube(5000, 1500.0, 4)
    zero = np.zeros(net.n, np.uint32)
    g, st = _update(_net(net.xyz, zero, net.vac_exc, net.vac_inh, net.signal, (net.box_min, net.box_max)), flags,
                    precision)
    assert st["searches"] == 0 and st["requests"] == 0
    g.close()
    g, st = _update(net, flags, precision, enabled=np.zeros(net.n, np.uint8))
    assert st["searches"] == 0 and st["requests"] == 0
    va, ve, vi = g.get_vacancies()
    assert np.array_equal(va, net.vac_axon) and np.array_equal(ve, net.vac_exc)
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_no_vacant_dendrites_every_search_fails(flags):
    # BarnesHutBase::find_target_neurons (BarnesHutBase.h:283-289): a failed search ends the source's
    # searches, so every source with vacant axons makes exactly one search and no request
    _say("no vacant dendrites", flags)
    net = uniform_cube(5000, 1500.0, 5)
    zero = np.zeros(net.n, np.uint32)
    g, st = _update(_net(net.xyz, net.vac_axon, zero, zero, net.signal, (net.box_min, net.box_max)), flags)
    assert st["searches"] == int((net.vac_axon > 0).sum())
    assert st["requests"] == 0 and st["created"] == 0
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_single_and_two_neurons(flags):
    _say("one and two neurons", flags)
    # one neuron: the root is a leaf holding the source itself, so no target (BarnesHutBase.h:231-237)
    g, st = _update(_net([[10.0, 20.0, 30.0]], [3], [2], [2]), flags)
    assert st["requests"] == 0
    g.close()
    # two neurons of opposite types: each picks the other while it has a vacant dendrite of its type
    g, st = _update(_net([[100.0, 100.0, 100.0], [900.0, 900.0, 900.0]], [2, 1], [5, 5], [5, 5], sig=[0, 1]), flags)
    t, s, sig, _ = g.get_requests(st["requests"])
    assert st["requests"] == 3 and np.all(t != s)
    assert st["created"] == 3
    va, ve, vi = g.get_vacancies()
    assert list(va) == [0, 0] and list(ve) == [5, 3] and list(vi) == [4, 5]
    g.close()


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
def test_neurons_on_box_faces_match_oracle(oracle, flags):
    # positions on the faces, edges and corners of the box, where the octant tests meet the boundary
    _say("box faces", flags)
    rng = np.random.default_rng(7)
    pts = rng.integers(0, 5, (3000, 3)).astype(np.float64) * 250.0   # a 5^3 grid including both faces
    pts = np.unique(pts, axis=0)
    jitter = rng.random((2000, 3)) * 1000.0
    xyz = np.vstack([pts, jitter])
    n = len(xyz)
    net = _net(xyz, rng.poisson(2, n), rng.poisson(1.5, n), rng.poisson(1.5, n), sig=(rng.random(n) < 0.2))
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    ref = t.find_targets(net.signal, net.vac_axon, seed=3, step=1, threads=0)
    t.close()
    g = BarnesHutHIP(net.box_min, net.box_max, seed=3, flags=flags)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.find_target_neurons(1)
    tgt, tests, cnt = g.get_search_slots(int(net.vac_axon.sum()))
    off = ref["offsets"].astype(np.int64)
    checked = 0
    for i in range(n):
        k = int(net.vac_axon[i])
        sl = slice(off[i], off[i] + k)
        if k == 0 or np.any(ref["margin"][sl] < 1e-12):
            continue
        c = int(cnt[i])
        assert c == int(ref["count"][i]), i
        assert np.array_equal(np.asarray(ref["target"])[sl][:c], tgt[sl][:c]), i
        e = min(c + 1, k)   # the tests of every executed search, a failed last one included
        assert np.array_equal(np.asarray(ref["node_tests"])[sl][:e], tests[sl][:e]), i
        checked += 1
    assert checked > 0.9 * int((net.vac_axon > 0).sum())
    g.close()


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_one_neuron_with_thousands_of_axons(oracle, precision):
    # 3000 vacant axons on one source: 375 rounds of 8 draws against one shared level-1 list
    _say("3000 axons on one source", precision)
    rng = np.random.default_rng(11)
    n = 4000
    xyz = rng.random((n, 3)) * 1000.0
    va = np.zeros(n, np.uint32)
    va[17] = 3000
    ve = np.ones(n, np.uint32)
    net = _net(xyz, va, ve, ve)
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    ref = t.find_targets(net.signal, net.vac_axon, seed=4, step=2, sigma=150.0, threads=0)
    t.close()
    g = BarnesHutHIP(net.box_min, net.box_max, seed=4, sigma=150.0, precision=precision)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    g.find_target_neurons(2)
    tgt, tests, cnt = g.get_search_slots(3000)
    assert int(cnt[17]) == int(ref["count"][17]) == 3000
    if precision == RL_PREC_FP64:
        clear = np.asarray(ref["margin"]) > 1e-12
        assert np.array_equal(np.asarray(ref["target"])[clear], tgt[clear])
    else:
        assert np.mean(np.asarray(ref["target"]) != tgt) < 1e-2
    assert np.all(tgt != 17)
    g.close()


def test_second_commit_of_one_search_is_refused():
    # the requests of one search are committed once: a second rl_bh_commit returns RL_ESTATE and leaves
    # the vacancies as the first left them (it would otherwise take every accepted dendrite and axon a
    # second time); the requests and accept flags stay readable, and a new search allows a commit again
    from relearn_amd._lib import BHError

    _say("second commit refused")
    net = uniform_cube(3000, 1000.0, 61)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=2)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    m = g.find_target_neurons(0)
    created = g.commit(0)
    assert m > 0 and created > 0
    vac = [v.copy() for v in g.get_vacancies()]
    with pytest.raises(BHError) as err:
        g.commit(0)
    assert err.value.code == -6
    assert all(np.array_equal(a, b) for a, b in zip(vac, g.get_vacancies()))
    assert len(g.get_requests(m)[0]) == m and int(g.get_commit_result(m).sum()) == created
    # a new search of the committed state can be committed
    g.prepare_resident()
    g.find_target_neurons(1)
    g.commit(1)
    # a replay with stats commits its own requests: a commit after it is refused too
    g.graph_create(0, int(net.vac_axon.sum()))
    g.graph_launch(0, 2, stats=True)
    with pytest.raises(BHError) as err:
        g.commit(2)
    assert err.value.code == -6
    g.close()
