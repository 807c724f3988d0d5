"""End-to-end parity of simulations with the oracle: 1000 neurons in a 1000 um cube (80/20 E/I) with
thresholds that make elements grow and shrink, the CPU reference configuration itself (1000 neurons,
20,000 steps, the reference's thresholds), and the reference's real 51,842-neuron positions with its
425,706-synapse input network.

Each period runs 100 neuron steps with background input (the same caller arrays on both sides) and
synaptic input over the network the updates build, the element commit of the three element types
with the deletion of the synapses of shrunken elements, and one connectivity update (leaf fill,
upward pass, search, commit).  The oracle restates each of these on the CPU (neuron step, synaptic
input, ElementBaseBase::update, SynapseDeletionFinderCPU, pointer octree search under the
counter-based draws, serial commit) and keeps its own network graph.  The calcium target lies inside
the initial calcium range, so some neurons grow elements while others shrink and delete synapses.
After every period the neuron state, the element counts, the vacancies, the created and deleted
synapses and the network graph must be identical.  FP64 only: the FP32-mixed picks are allowed to differ from FP64 (measured by bench.py).
"""
import os
import sys

import numpy as np
import pytest

from relearn_amd import RL_PREC_FP64, BarnesHutHIP, StepParams
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _pairs(off, src, w):
    t = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    out = {}
    for a, b, c in zip(t.tolist(), src.tolist(), w.tolist()):
        out[(a, b)] = out.get((a, b), 0) + c
    return {k: v for k, v in out.items() if v != 0}


def _simulate(oracle, xyz, signal, box_min, box_max, graph, periods, steps, seed, rng, cond=1.5,
              eta=(0.02, 0.01, 0.0), target=0.1, nu=0.02, ca0=(0.05, 0.15), start_elements=True, check_every=1,
              retract=(0.0, 0.0, 0.0), min_ca=None, retract_per_neuron=None, ca_init=None, target_ca=None,
              create=None):
    """Runs `periods` periods on the device and in the oracle side by side and compares them every
    `check_every` periods and after the last.

    retract: vacant retract ratio per element type; min_ca / retract_per_neuron: (3, n) per-neuron
    minimum calcium and retract ratios (rl_neurons_set_element_params); ca_init / target_ca: per-neuron
    initial and target calcium; create: (period, xyz, signal) of a creation event at the start of that
    period (rl_bh_create_neurons), the new neurons starting at x = c, u = 0, calcium 0, the target
    calcium `target` and no elements."""
    n = len(xyz)
    pg = StepParams.izhikevich(input_mode=0, nu=nu)
    pg.synaptic_input, pg.synapse_conductance = 1, cond
    po = oracle.StepParams()
    for f in ("a", "b", "c", "d", "v_spike", "k1", "k2", "k3", "h", "tau_ca", "beta_ca"):
        setattr(po, f, getattr(pg, f))
    for t in range(3):
        pg.eta[t] = po.eta[t] = eta[t]
        po.nu[t] = pg.nu[t]
        pg.retract_ratio[t] = retract[t]
    rr = np.repeat(np.asarray(retract, np.float64)[:, None], n, axis=1)
    if retract_per_neuron is not None:
        rr = np.array(retract_per_neuron, dtype=np.float64)
    mc = None if min_ca is None else np.array(min_ca, dtype=np.float64)
    og = oracle.Graph(n) if graph is None else oracle.Graph(n, *graph)
    # connected counts that match the initial graph, a fraction of an element vacant on top
    conn = np.zeros((3, n), np.uint32)
    eo, es, ew = og.export()
    t_of = np.repeat(np.arange(n), np.diff(eo.astype(np.int64)))
    np.add.at(conn[0], es, np.abs(ew))
    np.add.at(conn[1], t_of[ew > 0], ew[ew > 0])
    np.add.at(conn[2], t_of[ew < 0], -ew[ew < 0])
    if start_elements:
        grown = conn + (rng.poisson(1.0, (3, n)) if graph is None else rng.uniform(0.0, 1.5, (3, n)))
    else:
        grown = conn.copy()
    grown = grown.astype(np.float64)
    x = np.full(n, -65.0)
    u = np.full(n, -13.0)
    ca = rng.uniform(ca0[0], ca0[1], n) if ca0 is not None else np.zeros(n)
    if ca_init is not None:
        ca = np.array(ca_init, dtype=np.float64)
    tca = np.full(n, target) if target_ca is None else np.array(target_ca, dtype=np.float64)

    g = BarnesHutHIP(box_min, box_max, seed=seed, precision=RL_PREC_FP64)
    g.init(xyz, signal)
    g.net_init(len(es) + 50 * n, *(graph if graph is not None else (None, None, None)))
    g.neurons_init(pg, tca, x=x, u=u, ca=ca, grown=grown.ravel(), connected=conn.ravel())
    if mc is not None or retract_per_neuron is not None:
        g.neurons_set_element_params(mc, None if retract_per_neuron is None else rr)

    tree = oracle.Tree(xyz, box_min, box_max)
    xo, uo, co = x.copy(), u.copy(), ca.copy()
    fo = np.zeros(n, np.uint8)
    do = np.zeros(3 * n)
    go = grown.copy()
    vac = np.zeros((3, n), np.uint32)
    created_total = deleted_total = 0
    for per in range(periods):
        if create is not None and create[0] == per:
            # creation event (Neurons::create_neurons): both sides append the neurons; the oracle
            # rebuilds its tree and graph over all of them
            cx, cs = np.asarray(create[1], np.float64), np.asarray(create[2], np.uint8)
            k = len(cx)
            g.create_neurons(cx, cs, x=np.full(k, pg.c), u=np.zeros(k), ca=np.zeros(k), target_ca=np.full(k, target))
            xyz = np.concatenate([xyz, cx])
            signal = np.concatenate([signal, cs])
            tree.close()
            tree = oracle.Tree(xyz, box_min, box_max)
            eo, es, ew = og.export()
            og.close()
            og = oracle.Graph(n + k, np.concatenate([eo, np.full(k, eo[-1], eo.dtype)]), es, ew)
            xo = np.concatenate([xo, np.full(k, pg.c)])
            uo = np.concatenate([uo, np.zeros(k)])
            co = np.concatenate([co, np.zeros(k)])
            fo = np.concatenate([fo, np.zeros(k, np.uint8)])
            tca = np.concatenate([tca, np.full(k, target)])
            do = np.concatenate([do.reshape(3, n), np.zeros((3, k))], axis=1).ravel()
            go = np.concatenate([go, np.zeros((3, k))], axis=1)
            conn = np.concatenate([conn, np.zeros((3, k), np.uint32)], axis=1)
            vac = np.concatenate([vac, np.zeros((3, k), np.uint32)], axis=1)
            rr = np.concatenate([rr, np.repeat(np.asarray(retract, np.float64)[:, None], k, axis=1)], axis=1)
            if mc is not None:
                mc = np.concatenate([mc, np.repeat(np.asarray(eta, np.float64)[:, None], k, axis=1)], axis=1)
            n += k
            _say("created", k, "neurons before period", per, "now", n)
        off, src, w = og.export()
        for k in range(steps):
            bg = rng.normal(5.0, 2.0, n)
            inp = oracle.synaptic_input(off, src, w, fo, cond, bg)
            oracle.neuron_step(po, xo, uo, co, inp, fo, tca, do, None, min_calcium=mc)
            g.neurons_step(per * steps + k, bg)
        dv = do.reshape(3, n)
        dref = np.zeros((3, n), np.uint32)
        for t in range(3):
            gt, dt, ct, vt = go[t].copy(), dv[t].copy(), conn[t].copy(), vac[t].copy()
            dref[t] = oracle.elements_update(gt, dt, ct, vt, retract_ratio=rr[t].copy())
            go[t], dv[t], conn[t], vac[t] = gt, dt, ct, vt
            deleted_total += og.delete(t, dref[t], signal, conn, vac, step=per, seed=seed)
        do = dv.ravel().copy()
        assert np.array_equal(g.neurons_commit_elements(per), dref), per
        # connectivity update in the oracle: search under the counter-based draws, serial commit
        tree.update(vac[1], vac[2])
        r = tree.find_targets(signal, vac[0], seed=seed, step=per)
        assert np.all(r["margin"] > 1e-12), "a draw within 1e-12 of a boundary: picks may legitimately differ"
        cnt = r["count"].astype(np.int64)
        rs = np.repeat(np.arange(n), cnt)
        rj = np.concatenate([np.arange(c) for c in cnt]) if cnt.sum() else np.zeros(0, np.int64)
        rt = r["target"][r["offsets"].astype(np.int64)[rs] + rj]
        rg = signal[rs]
        ve, vi, va = vac[1].copy(), vac[2].copy(), vac[0].copy()
        ce, ci, cax = conn[1].copy(), conn[2].copy(), conn[0].copy()
        ok, created = oracle.commit(rt, rs, rg, rj, ve, vi, va, seed=seed, step=per, conn_exc=ce, conn_inh=ci,
                                    conn_axon=cax)
        vac[1], vac[2], vac[0], conn[1], conn[2], conn[0] = ve, vi, va, ce, ci, cax
        acc = ok.astype(bool)
        og.add(rt[acc], rs[acc], np.where(rg[acc] != 0, -1, 1))
        st = g.update_connectivity(per)
        created_total += created
        assert st["requests"] == len(rt) and st["created"] == created, per
        if (per + 1) % check_every != 0 and per + 1 != periods:
            continue
        _say("period", per, "requests", len(rt), "created", created, "deleted so far", deleted_total)
        # compare the whole state
        sg = g.neurons_get()
        for key, ref in (("x", xo), ("u", uo), ("ca", co), ("fired", fo)):
            assert np.array_equal(sg[key], ref), (per, key)
        assert np.allclose(sg["delta"].ravel(), do, rtol=1e-9, atol=1e-300), per
        assert np.allclose(sg["grown"], go, rtol=1e-9, atol=1e-12), per
        assert np.array_equal(sg["connected"], conn), per
        gv = g.get_vacancies()
        for t in range(3):
            assert np.array_equal(gv[t], vac[t]), (per, t)
        assert _pairs(*g.net_get_edges()) == _pairs(*og.export()), per
    tree.close()
    og.close()
    g.close()
    return created_total, deleted_total


def test_small_simulation_matches_oracle(oracle):
    n, periods, steps, seed = 1000, 6, 100, 5
    _say("simulation parity", n, "neurons,", periods, "periods of", steps, "steps")
    net = uniform_cube(n, 1000.0, 0)
    created, deleted = _simulate(oracle, net.xyz, net.signal, net.box_min, net.box_max, None, periods, steps, seed,
                                 np.random.default_rng(17))
    assert created > 0 and deleted > 0


def test_reference_cpu_configuration_simulation_matches_oracle(oracle):
    # the CPU reference configuration: 1,000 neurons uniform in a 1000 um cube (seed 0, 80/20 E/I),
    # Izhikevich defaults, Normal(5, 2) input, calcium tau 1e4 and beta 1e-3, growth rate 1e-4 and the
    # reference's thresholds (eta 0.4 / 0.1 / 0.0, target calcium 0.7), elements and calcium from 0,
    # synaptic input at conductance 0.03, plasticity every 100 steps, 20,000 steps; the state must
    # stay identical for the whole run (calcium stays below the axons' eta, so no axon grows and no
    # synapse forms: the reference's own outcome for this configuration)
    n, periods, steps, seed = 1000, 200, 100, 0
    _say("1000-neuron, 20,000-step simulation parity", n, "neurons,", periods, "periods of", steps, "steps")
    net = uniform_cube(n, 1000.0, 0)
    _simulate(oracle, net.xyz, net.signal, net.box_min, net.box_max, None, periods, steps, seed,
              np.random.default_rng(19), cond=0.03, eta=(0.4, 0.1, 0.0), target=0.7, nu=1e-4, ca0=None,
              start_elements=False, check_every=20)


def test_real_network_simulation_matches_oracle(oracle):
    # the reference's input/positions.txt (51,842 neurons) with its input/rank_0_in_network.txt
    # (425,706 synapses), read from the committed fixtures
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    e = np.load(os.path.join(GOLDEN, "network_51842.npz"))
    n = len(d["xyz"])
    order = np.argsort(e["target"], kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(e["target"], minlength=n))]).astype(np.uint64)
    graph = (off, e["source"][order].astype(np.uint32), np.ones(len(order), np.int32))
    _say("real network simulation parity", n, "neurons,", len(order), "synapses")
    created, deleted = _simulate(oracle, d["xyz"], d["signal"], d["box_min"], d["box_max"], graph, 2, 20, 8,
                                 np.random.default_rng(18), cond=0.03)
    assert deleted > 0


def test_simulation_with_retract_ratio_matches_oracle(oracle):
    # --retract-ratio (relearn.cpp:617-618): a fraction of the vacant elements retracts at every element
    # commit (ElementBaseBase::update, ElementBaseBase.h:420-490), one ratio per element type
    n, periods, steps, seed = 1000, 4, 100, 21
    _say("simulation parity with retract ratios 0.1 / 0.25 / 0.5")
    net = uniform_cube(n, 1000.0, 2)
    created, deleted = _simulate(oracle, net.xyz, net.signal, net.box_min, net.box_max, None, periods, steps, seed,
                                 np.random.default_rng(23), retract=(0.1, 0.25, 0.5))
    assert created > 0


def test_simulation_with_creation_event_matches_oracle(oracle):
    # Neurons::create_neurons (Simulation.cpp:425) between periods: 200 neurons join a 1000-neuron
    # network; the octree, the searches, the element state and the network graph must go on matching
    n, periods, steps, seed = 1000, 4, 100, 25
    _say("simulation parity with a creation event of 200 neurons")
    net = uniform_cube(n, 1000.0, 4)
    rng = np.random.default_rng(26)
    cx = rng.random((200, 3)) * 1000.0
    cs = (rng.random(200) < 0.2).astype(np.uint8)
    created, deleted = _simulate(oracle, net.xyz, net.signal, net.box_min, net.box_max, None, periods, steps, seed,
                                 np.random.default_rng(27), create=(2, cx, cs))
    assert created > 0


def test_reference_text_files_drive_a_simulation(oracle, tmp_path):
    # The reference's own input files (the header and first 200 neurons of input/positions.txt, the
    # first 40 synapses of input/rank_0_in_network.txt) and a calcium file and a synaptic-elements file
    # in the layouts of CalciumIO / SynapticElementsIO, all read through relearn_amd.io into rl_bh_init,
    # rl_net_init, rl_neurons_init and rl_neurons_set_element_params; three periods against the oracle.
    from relearn_amd import io

    pos = io.read_neurons(os.path.join(GOLDEN, "positions_head.txt"))
    n = len(pos["xyz"])
    # the head of the in-network file names a few sources beyond the first n neurons: keep the synapses
    # whose two ends are both among them, as a rank that loads this subset would
    with open(os.path.join(GOLDEN, "network_head_in.txt")) as fh:
        kept = [ln for ln in fh if ln.startswith("#") or max(int(ln.split()[1]), int(ln.split()[3])) <= n]
    (tmp_path / "in_network.txt").write_text("".join(kept))
    off, src, w, _ = io.read_in_synapses(tmp_path / "in_network.txt", n)
    assert len(src) >= 30
    rng = np.random.default_rng(31)
    groups = io.groups_from_areas(pos["area"])
    assert "lh_middletemporal" in groups
    lines = ["# <description> <initial calcium> <target calcium>", "default 0.05 0.1", "lh_middletemporal 0.12 0.08"]
    listed = [i for i in range(n) if i % 3 == 0 and i not in groups["lh_middletemporal"]]
    for i in listed:
        lines.append(f"0:{i + 1} {rng.uniform(0.05, 0.15):.6f} {rng.uniform(0.08, 0.12):.6f}")
    (tmp_path / "calcium.txt").write_text("\n".join(lines) + "\n")
    ca_init, ca_tgt = io.read_calcium(tmp_path / "calcium.txt", n, groups)
    assert np.all(ca_init[groups["lh_middletemporal"]] == 0.12)
    params = {f: np.zeros((3, n)) for f in io.ELEMENT_FIELDS}
    params["min_calcium"][0] = rng.uniform(0.01, 0.04, n)
    params["min_calcium"][1] = rng.uniform(0.005, 0.02, n)
    params["retract_ratio"][:] = rng.uniform(0.0, 0.3, (3, n))
    params["nu"][:] = 0.02
    params["max_elements"][:] = 2.0
    io.write_synaptic_elements(tmp_path / "elements.txt", params)
    se = io.read_synaptic_elements(tmp_path / "elements.txt", n)
    _say("reference text files: 200 neurons, 40 synapses, per-neuron calcium and element parameters")
    created, deleted = _simulate(oracle, pos["xyz"], pos["signal"], pos["box_min"], pos["box_max"], (off, src, w), 3,
                                 100, 33, rng, ca_init=ca_init, target_ca=ca_tgt, min_ca=se["min_calcium"],
                                 retract_per_neuron=se["retract_ratio"])
    assert created > 0


def test_reference_configuration_with_growing_thresholds_matches_oracle(oracle):
    # the shape of the CPU reference configuration (1,000 neurons, 200 plasticity periods of 100 steps,
    # 20,000 steps, elements and calcium from 0, synaptic input at conductance 0.03) with thresholds at
    # which calcium lets elements grow (eta 0.02 / 0.01 / 0.0, target calcium 0.2, growth rate 1e-3):
    # synapses form through the whole run, so every connectivity update (search, commit, network
    # append, synaptic input over the grown network) is checked against the oracle, not only the step
    n, periods, steps, seed = 1000, 200, 100, 0
    _say("1000-neuron, 20,000-step simulation parity with growing thresholds")
    net = uniform_cube(n, 1000.0, 0)
    created, deleted = _simulate(oracle, net.xyz, net.signal, net.box_min, net.box_max, None, periods, steps, seed,
                                 np.random.default_rng(23), cond=0.03, eta=(0.02, 0.01, 0.0), target=0.2, nu=1e-3,
                                 ca0=None, start_elements=False, check_every=20)
    _say("  created", created, "deleted", deleted)
    assert created > 1000
