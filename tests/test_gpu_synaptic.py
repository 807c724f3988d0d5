"""Network graph and synaptic input on the device (rl_net_init, rl_step_params.synaptic_input).

Bar: with synaptic input the neuron step stays bit-exact against the oracle (x, u, Ca, fired; deltas
within 1e-9 relative), the device graph after a commit holds exactly the initial edges plus one unit
of weight per accepted request (+1 for an excitatory source, -1 for an inhibitory one), one entry
per (source, target) pair, and a captured
period with synaptic input and graph append reproduces the eager calls bit for bit.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _random_graph(n, m, seed, wmax=2):
    rng = np.random.default_rng(seed)
    tgt = rng.integers(0, n, m)
    src = rng.integers(0, n, m)
    w = rng.integers(-wmax, wmax + 1, m).astype(np.int32)
    order = np.argsort(tgt, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(tgt, minlength=n))]).astype(np.uint64)
    return off, src[order].astype(np.uint32), w[order]


def _pairs(off, src, w):
    """(target, source) -> summed weight, zero pairs dropped: the graph as NetworkGraph keeps it."""
    t = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    out = {}
    for a, b, c in zip(t.tolist(), src.tolist(), w.tolist()):
        out[(a, b)] = out.get((a, b), 0) + c
    return {k: v for k, v in out.items() if v != 0}


def _merged(off, src, w):
    """Device graphs hold one entry per pair: check that, and return the pairs."""
    t = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    assert len(set(zip(t.tolist(), src.tolist()))) == len(src), "a (source, target) pair appears twice"
    return _pairs(off, src, w)


def _izh_params(oracle_mod, synaptic, conductance):
    pg = StepParams.izhikevich(input_mode=0)
    pg.synaptic_input = 1 if synaptic else 0
    pg.synapse_conductance = conductance
    po = oracle_mod.StepParams()
    for f in ("a", "b", "c", "d", "v_spike", "k1", "k2", "k3", "h", "tau_ca", "beta_ca"):
        setattr(po, f, getattr(pg, f))
    for k in range(3):
        po.eta[k] = pg.eta[k]
        po.nu[k] = pg.nu[k]
    return pg, po


def test_synaptic_step_parity(oracle):
    n, steps = 3000, 40
    _say("synaptic step parity", n, "neurons", steps, "steps")
    rng = np.random.default_rng(3)
    off, src, w = _random_graph(n, 12 * n, 4)
    pg, po = _izh_params(oracle, True, 3.0)
    x = rng.uniform(-70, -50, n)
    u = rng.uniform(-15, -10, n)
    ca = rng.uniform(0, 0.3, n)
    tca = np.full(n, 0.7)
    en = np.ones(n, np.uint8)
    en[::11] = 0
    g = BarnesHutHIP((0, 0, 0), (1, 1, 1))
    g.init(rng.random((n, 3)), np.zeros(n, np.uint8))
    g.set_enabled(en)
    g.net_init(len(src) + 100, off, src, w)
    g.neurons_init(pg, tca, x=x, u=u, ca=ca)
    xo, uo, co = x.copy(), u.copy(), ca.copy()
    fo = np.zeros(n, np.uint8)
    do = np.zeros(3 * n)
    fired_any = 0
    bgs = []
    for s in range(steps):
        bg = rng.normal(5.0, 2.0, n)
        bgs.append(bg)
        inp = oracle.synaptic_input(off, src, w, fo, 3.0, bg)
        oracle.neuron_step(po, xo, uo, co, inp, fo, tca, do, en)
        g.neurons_step(s, bg)
        fired_any += int(fo.sum())
    st = g.neurons_get()
    assert fired_any > 0
    assert np.array_equal(st["x"], xo) and np.array_equal(st["u"], uo) and np.array_equal(st["ca"], co)
    assert np.array_equal(st["fired"], fo)
    assert np.allclose(st["delta"].ravel(), do, rtol=1e-9, atol=1e-300)
    # the synaptic input matters: the oracle run with the background input alone ends elsewhere
    xb, ub, cb = x.copy(), u.copy(), ca.copy()
    fb, db = np.zeros(n, np.uint8), np.zeros(3 * n)
    for bg in bgs:
        oracle.neuron_step(po, xb, ub, cb, bg, fb, tca, db, en)
    assert not np.array_equal(xb, xo)
    g.close()


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_net_append_matches_commit(precision):
    _say("network graph append after commits, precision", precision)
    net = uniform_cube(20_000, 3000.0, 41)
    n = net.n
    off0, src0, w0 = _random_graph(n, 5 * n, 5, wmax=1)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=9, precision=precision)
    g.init(net.xyz, net.signal)
    g.net_init(len(src0) + 1000, off0, src0, w0)  # small capacity: the eager append grows it
    expect = [(np.repeat(np.arange(n), np.diff(off0.astype(np.int64))), src0.astype(np.int64), w0.astype(np.int64))]
    rng = np.random.default_rng(6)
    for upd in range(3):
        va = rng.poisson(2.0, n).astype(np.uint32)
        ve = rng.poisson(1.5, n).astype(np.uint32)
        vi = rng.poisson(1.5, n).astype(np.uint32)
        g.prepare_update_connectivity(va, ve, vi)
        st = g.update_connectivity(upd)
        t, s, sig, _ = g.get_requests(st["requests"])
        ok = g.get_commit_result(st["requests"]).astype(bool)
        assert ok.sum() == st["created"] > 0
        expect.append((t[ok].astype(np.int64), s[ok].astype(np.int64), np.where(sig[ok] != 0, -1, 1)))
    et = np.concatenate([e[0] for e in expect])
    es = np.concatenate([e[1] for e in expect])
    ew = np.concatenate([e[2] for e in expect])
    order = np.argsort(et, kind="stable")
    eoff = np.concatenate([[0], np.cumsum(np.bincount(et, minlength=n))]).astype(np.uint64)
    assert _merged(*g.net_get_edges()) == _pairs(eoff, es[order], ew[order])
    g.close()


@pytest.mark.parametrize("room", [20, 0])
def test_graph_period_with_synaptic_input_matches_eager(room):
    """room: free network graph entries per neuron at the start; with none, the first replays' synapses
    do not fit, and the replay grows the graph, appends them and captures the period again."""
    n, steps, periods = 10_000, 21, 3   # an odd step count exercises the fired-status copy-back
    _say("graph period with synaptic input vs eager,", n, "neurons,", steps, "steps x", periods, "room", room)
    net = uniform_cube(n, 2200.0, 42)
    off0, src0, w0 = _random_graph(n, 8 * n, 7, wmax=1)
    pairs0 = len(_pairs(off0, src0, w0))   # the entries the graph starts with
    ctxs = []
    for _ in range(2):
        c = BarnesHutHIP(net.box_min, net.box_max, seed=13, precision=RL_PREC_FP32_MIXED)
        c.init(net.xyz, net.signal)
        c.net_init(pairs0 + room * n, off0, src0, w0)
        p = StepParams.izhikevich(input_mode=1, nu=0.02)
        p.synaptic_input = 1
        p.synapse_conductance = 2.0
        for t in range(3):
            p.eta[t] = (0.02, 0.01, 0.0)[t]
        rng = np.random.default_rng(8)
        c.neurons_init(p, np.full(n, 0.2), x=np.full(n, -65.0), u=np.full(n, -13.0),
                       grown=rng.poisson(1.0, (3, n)).astype(np.float64).ravel())
        ctxs.append(c)
    eager, graph = ctxs
    graph.neurons_step(0)       # leaves the last status in the second buffer before capture
    eager.neurons_step(0)
    graph.graph_create(steps, 8 * n)
    regrown = 0
    for per in range(periods):
        first = 1 + per * steps
        for k in range(steps):
            eager.neurons_step(first + k)
        eager.neurons_commit_elements()
        se_stats = eager.update_connectivity(per)
        sg_stats = graph.graph_launch(first, per)
        _say("period", per, "created", sg_stats["created"], "net regrown", sg_stats["net_regrown"])
        regrown += sg_stats["net_regrown"]
        assert sg_stats["net_capacity"] >= len(graph.net_get_edges()[1])
        assert se_stats["created"] == sg_stats["created"]
        se, sg = eager.neurons_get(), graph.neurons_get()
        for key in ("x", "u", "ca", "fired", "delta", "grown", "connected"):
            assert np.array_equal(se[key], sg[key]), (per, key)
        assert _merged(*eager.net_get_edges()) == _merged(*graph.net_get_edges())
    assert len(_pairs(*eager.net_get_edges())) > len(_pairs(off0, src0, w0))
    assert (regrown > 0) == (room == 0)
    eager.close()
    graph.close()


def test_synaptic_input_needs_graph():
    net = uniform_cube(100, 100.0, 43)
    g = BarnesHutHIP(net.box_min, net.box_max)
    g.init(net.xyz, net.signal)
    p = StepParams.izhikevich(input_mode=1)
    p.synaptic_input = 1
    g.neurons_init(p, np.full(100, 0.7))
    with pytest.raises(BHError):
        g.neurons_step(0)
    with pytest.raises(BHError):
        g.net_init(10, np.array([0, 5] + [5] * 99, np.uint64), np.zeros(5, np.uint32) + 200, np.ones(5, np.int32))
    g.net_init(10)
    g.neurons_step(0)
    g.close()
