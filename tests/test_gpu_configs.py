"""The batched and the full-simulation workloads at their stated sizes.

- The 256-network parameter sweep (sizes log-uniform in [2,000, 200,000], each network uniform in a
  cube of side 1000 (n/1000)^(1/3) um, 80/20 E/I, Poisson vacancies, theta 0.3, sigma 750 um) in one
  context: one connectivity update with per-network invariants, three of its networks (the smallest,
  the median and the largest) checked directly against the CPU oracle with their own draw-stream
  seeds, and FP32-mixed picks within 1e-4 of FP64; then 10^4 steps (100 captured periods of 100
  steps) with growth rates drawn per network from [5e-5, 5e-4].
- 4M neurons in a 16,000 um cube: 10^5 steps (1000 captured periods of 100 steps), FP32-mixed next
  to FP64, calcium and synapse trajectories compared at 10k, 50k and 100k steps.

Both simulations run with the reference's parameters as stated (growth thresholds eta 0.4 / 0.1 / 0.0
for axons / excitatory / inhibitory dendrites, target calcium 0.7) and with thresholds that let
axons grow (eta 0.02 / 0.01 / 0.0, target 0.2).  The oracle cannot run these sizes in seconds; the
checks are size-independent properties after the last period:

- every neuron's connected + vacant elements of each type equal floor(grown) (ElementBaseBase::update,
  ElementBaseBase.h:420-490, with the synapse commit and the partner disconnections moving elements
  between the two);
- the connected axons, the connected dendrites and the network graph's synapses are one count;
- synapses join two different neurons of one network and carry the source's sign;
- FP32-mixed and FP64 trajectories agree (bit for bit when no synapse forms: the neuron step is FP64
  in both modes and the search then has nothing to do).
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle
from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams
from relearn_amd._lib import network_seed
from relearn_amd.workloads import parameter_sweep, uniform_cube

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF
REFERENCE = {"eta": (0.4, 0.1, 0.0), "target": 0.7}
GROWING = {"eta": (0.02, 0.01, 0.0), "target": 0.2}


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sweep():
    return parameter_sweep(256, seed=0)


def _bounds(ens):
    return ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0)


def test_parameter_sweep_update_256_networks(sweep):
    ens = sweep
    n, G = ens.n, ens.num_networks
    seed, step = 5, 2
    off = ens.offsets.astype(np.int64)
    net_of = np.repeat(np.arange(G), np.diff(off))
    va = ens.vac_axon.astype(np.int64)
    slot_off = np.concatenate([[0], np.cumsum(va)])
    sn = np.repeat(np.arange(n), va)
    jj = np.arange(int(va.sum())) - slot_off[sn]
    picks = {}
    for prec in (RL_PREC_FP64, RL_PREC_FP32_MIXED):
        _say(f"parameter sweep: {G} networks, {n} neurons, one update, precision {prec}")
        g = BarnesHutHIP(*_bounds(ens), seed=seed, precision=prec)
        g.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
        g.prepare_update_connectivity(ens.vac_axon, ens.vac_exc, ens.vac_inh)
        nreq = g.find_target_neurons(step)
        tgt, tests, cnt = g.get_search_slots(int(va.sum()))
        c = cnt.astype(np.int64)
        assert np.all(c <= va) and int(c.sum()) == nreq
        found = jj < c[sn]
        assert np.all(tgt[found] != NONE) and np.all(tgt[jj > c[sn]] == NONE)
        picked, src = tgt[found].astype(np.int64), sn[found]
        assert np.all(picked != src)
        assert np.array_equal(net_of[picked], net_of[src]), "a pick left its network"
        exc = ens.signal[src] == 0
        assert np.all(ens.vac_exc[picked[exc]] > 0) and np.all(ens.vac_inh[picked[~exc]] > 0)
        created = g.commit(step)
        t, s, sig, _ = g.get_requests(nreq)
        ok = g.get_commit_result(nreq).astype(bool)
        assert int(ok.sum()) == created > 0
        vax, vex, vin = g.get_vacancies()
        assert np.array_equal(vex.astype(np.int64), ens.vac_exc - np.bincount(t[ok & (sig == 0)], minlength=n))
        assert np.array_equal(vin.astype(np.int64), ens.vac_inh - np.bincount(t[ok & (sig != 0)], minlength=n))
        assert np.array_equal(vax.astype(np.int64), va - np.bincount(s[ok], minlength=n))
        g.close()
        picks[prec] = (tgt, tests, cnt)
        _say(f"  {nreq} requests, {created} synapses")
    # three networks against the oracle, each with its own draw stream (rl_bh_network_seed)
    sizes = np.diff(off)
    order = np.argsort(sizes, kind="stable")
    tgt, tests, cnt = picks[RL_PREC_FP64]
    near_total = 0
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        a, b = int(off[k]), int(off[k + 1])
        _say(f"  network {k} ({b - a} neurons) against the oracle")
        t = pyoracle.Tree(ens.xyz[a:b], ens.boxes[k, :3], ens.boxes[k, 3:])
        t.update(ens.vac_exc[a:b], ens.vac_inh[a:b], threads=16)
        ref = t.find_targets(ens.signal[a:b], ens.vac_axon[a:b], seed=network_seed(seed, k), step=step, threads=16)
        t.close()
        sa, sb = int(slot_off[a]), int(slot_off[b])
        mine = tgt[sa:sb].astype(np.int64)
        mine = np.where(mine == NONE, NONE, mine - a)
        assert np.array_equal(cnt[a:b], ref["count"])
        near = ref["margin"] < 1e-12
        assert np.all((mine == ref["target"]) | near)
        searched = ref["target"] != NONE
        assert np.array_equal(tests[sa:sb][searched], ref["node_tests"][searched])
        near_total += int(near.sum())
    _say(f"  oracle parity on three networks, {near_total} near-boundary draws")
    a, b = picks[RL_PREC_FP64][0], picks[RL_PREC_FP32_MIXED][0]
    div = float(np.mean(a != b))
    _say(f"  FP32-mixed vs FP64 pick divergence {div:.3g}")
    assert div <= 1e-4


def _simulate(make, n, periods, period, params, checkpoints, nu_per_network=None, nu=1e-4):
    """Captured periods from zero elements; returns (trajectory rows at the checkpoint steps, final state)."""
    g = make()
    p = StepParams.izhikevich(input_mode=1, nu=nu)
    for t in range(3):
        p.eta[t] = params["eta"][t]
    p.synaptic_input = 1
    p.synapse_conductance = 0.03
    zero = np.zeros(n, np.uint32)
    g.prepare_update_connectivity(zero, zero, zero)
    g.net_init(8 * n)
    g.neurons_init(p, np.full(n, params["target"]), x=np.full(n, -65.0), u=np.full(n, -13.0))
    if nu_per_network is not None:
        g.neurons_set_network_growth(np.repeat(nu_per_network, 3))
    cap = max(n // 2, 1024)
    g.graph_create(period, cap)
    rows, created, deleted = [], 0, 0
    for per in range(periods):
        try:
            st = g.graph_launch(per * period, per)
        except BHError as e:
            if e.code != -7 or "max_vacant_axons" not in str(e):
                raise
            st = g.update_connectivity(per)
            cap = max(2 * cap, int(g.get_vacancies()[0].sum()) * 5 // 4)
            g.graph_create(period, cap)
        created += st["created"]
        deleted += st.get("deleted", 0)
        if (per + 1) * period in checkpoints:
            s = g.neurons_get()
            rows.append((float(s["ca"].mean()), int(s["connected"][0].sum())))
    s = g.neurons_get()
    vac = g.get_vacancies()
    off, src, w = g.net_get_edges()
    g.close()
    return rows, {"state": s, "vac": vac, "edges": (off, src, w), "created": created, "deleted": deleted}


def _check_final(f, signal, net_of):
    s, vac = f["state"], f["vac"]
    con = s["connected"].astype(np.int64)
    for t in range(3):
        assert np.array_equal(con[t] + vac[t].astype(np.int64), np.floor(s["grown"][t]).astype(np.int64)), t
    off, src, w = f["edges"]
    total = int(np.abs(w.astype(np.int64)).sum())
    assert int(con[0].sum()) == int(con[1].sum() + con[2].sum()) == total
    if len(src) and total:
        tgt = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
        src = src.astype(np.int64)
        assert np.all(tgt != src)
        assert np.array_equal(net_of[tgt], net_of[src])
        assert np.all((w > 0) == (signal[src] == 0))
    return total


@pytest.mark.parametrize("params", ["reference", "growing"])
def test_parameter_sweep_simulation_10k_steps(sweep, params):
    ens = sweep
    n, G = ens.n, ens.num_networks
    par = REFERENCE if params == "reference" else GROWING
    net_of = np.repeat(np.arange(G), np.diff(ens.offsets.astype(np.int64)))
    checkpoints = (5000, 10000)
    res = {}
    for prec in (RL_PREC_FP32_MIXED, RL_PREC_FP64):
        _say(f"parameter sweep simulation, {params} thresholds, precision {prec}: 100 periods of 100 steps")

        def make():
            g = BarnesHutHIP(*_bounds(ens), seed=7, precision=prec)
            g.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
            return g

        rows, fin = _simulate(make, n, 100, 100, par, checkpoints, nu_per_network=ens.nu)
        syn = _check_final(fin, ens.signal, net_of)
        _say(f"  mean calcium {rows[-1][0]:.6f}, {syn} synapses, {fin['created']} created, {fin['deleted']} deleted")
        res[prec] = (rows, syn)
    a, b = res[RL_PREC_FP64], res[RL_PREC_FP32_MIXED]
    if params == "reference":
        # calcium stays below the axons' threshold 0.4: no axon grows, no synapse forms
        assert a[1] == b[1] == 0
        assert a[0] == b[0]
    else:
        assert a[1] > 0
        for (ca64, s64), (ca32, s32) in zip(a[0], b[0]):
            assert abs(ca32 - ca64) <= 1e-4 * abs(ca64)
            assert abs(s32 - s64) <= 1e-3 * max(s64, 1)


@pytest.mark.parametrize("params", ["reference", "growing"])
def test_4m_simulation_100k_steps(params):
    n, side = 4_000_000, 16_000.0
    par = REFERENCE if params == "reference" else GROWING
    net = uniform_cube(n, side, seed=3)
    net_of = np.zeros(n, np.int64)
    checkpoints = (10_000, 50_000, 100_000)
    res = {}
    for prec in (RL_PREC_FP32_MIXED, RL_PREC_FP64):
        _say(f"4M simulation, {params} thresholds, precision {prec}: 1000 periods of 100 steps")

        def make():
            g = BarnesHutHIP(net.box_min, net.box_max, seed=9, precision=prec)
            g.init(net.xyz, net.signal)
            return g

        rows, fin = _simulate(make, n, 1000, 100, par, checkpoints)
        syn = _check_final(fin, net.signal, net_of)
        _say(f"  trajectory {rows}, {syn} synapses")
        res[prec] = (rows, syn)
    a, b = res[RL_PREC_FP64], res[RL_PREC_FP32_MIXED]
    if params == "reference":
        assert a[1] == b[1] == 0
        assert a[0] == b[0]
    else:
        assert a[1] > n
        for (ca64, s64), (ca32, s32) in zip(a[0], b[0]):
            assert abs(ca32 - ca64) <= 1e-4 * abs(ca64)
            assert abs(s32 - s64) <= 1e-3 * max(s64, 1)
