"""Ensembles of independent networks in one context (rl_bh_init_networks).

Network g of an ensemble must reproduce, bit for bit, a one-network context created with the seed
rl_bh_network_seed(seed, g) on the same neurons: the same octree, searches, requests, commit and
neuron-step results.  The one-network path is itself checked against the CPU oracle in
test_gpu_parity.py, so these tests carry that parity over to ensembles.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_FLAG_EXACT_ORDER, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams
from relearn_amd._lib import network_seed

NONE = 0xFFFFFFFF


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _ensemble(sizes, seed):
    """Networks of the given sizes, each uniform in its own cube placed at its own offset."""
    rng = np.random.default_rng(seed)
    xyz, sig, boxes = [], [], []
    for g, n in enumerate(sizes):
        side = 1000.0 * max(n / 1000.0, 0.05) ** (1 / 3)
        lo = np.array([37.0 * g, -11.0 * g, 5.0 * g])
        boxes.append(np.concatenate([lo, lo + side]))
        xyz.append(lo + rng.random((n, 3)) * side)
        s = np.zeros(n, np.uint8)
        s[rng.permutation(n)[int(np.ceil(0.8 * n)):]] = 1
        sig.append(s)
    n = sum(sizes)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    va = rng.poisson(2.0, n).astype(np.uint32)
    ve = rng.poisson(1.5, n).astype(np.uint32)
    vi = rng.poisson(1.5, n).astype(np.uint32)
    return off, np.array(boxes), np.concatenate(xyz), np.concatenate(sig), va, ve, vi


def _run(ctx, xyz, sig, va, ve, vi, step, init):
    init(ctx)
    ctx.prepare_update_connectivity(va, ve, vi)
    nreq = ctx.find_target_neurons(step)
    slots = int(va.astype(np.uint64).sum())
    tgt, tests, cnt = ctx.get_search_slots(slots)
    created = ctx.commit(step)
    t, s, g, j = ctx.get_requests(nreq)
    ok = ctx.get_commit_result(nreq)
    vac = ctx.get_vacancies()
    return dict(tgt=tgt, tests=tests, cnt=cnt, created=created, t=t, s=s, g=g, j=j, ok=ok, vac=vac)


@pytest.mark.parametrize("precision,flags", [(RL_PREC_FP64, 0), (RL_PREC_FP32_MIXED, 0),
                                             (RL_PREC_FP32_MIXED, RL_FLAG_EXACT_ORDER)])
def test_ensemble_matches_single_networks(precision, flags):
    sizes = [1, 2, 700, 3000, 12000, 5]
    _say("ensemble vs single networks", sizes, "precision", precision, "flags", flags)
    off, boxes, xyz, sig, va, ve, vi = _ensemble(sizes, 41)
    seed, step = 123, 4
    big = BarnesHutHIP(boxes[:, :3].min(0), boxes[:, 3:].max(0), seed=seed, precision=precision, flags=flags)
    E = _run(big, xyz, sig, va, ve, vi, step, lambda c: c.init_networks(off, boxes, xyz, sig))
    m_all, _ = big.tree_info()
    big.close()
    slot_off = np.concatenate([[0], np.cumsum(va.astype(np.uint64))]).astype(np.int64)
    req_base = 0
    m_sum = 0
    for g in range(len(sizes)):
        a, b = int(off[g]), int(off[g + 1])
        one = BarnesHutHIP(boxes[g, :3], boxes[g, 3:], seed=network_seed(seed, g), precision=precision, flags=flags)
        S = _run(one, xyz[a:b], sig[a:b], va[a:b], ve[a:b], vi[a:b], step,
                 lambda c: c.init(xyz[a:b], sig[a:b]))
        m_g, _ = one.tree_info()
        m_sum += m_g
        one.close()
        sa, sb = slot_off[a], slot_off[b]
        et = E["tgt"][sa:sb].astype(np.int64)
        et = np.where(et == NONE, NONE, et - a)
        assert np.array_equal(et, S["tgt"].astype(np.int64)), g
        assert np.array_equal(E["tests"][sa:sb], S["tests"]), g
        assert np.array_equal(E["cnt"][a:b], S["cnt"]), g
        k = len(S["t"])
        assert np.array_equal(E["t"][req_base:req_base + k].astype(np.int64) - a, S["t"].astype(np.int64)), g
        assert np.array_equal(E["s"][req_base:req_base + k].astype(np.int64) - a, S["s"].astype(np.int64)), g
        assert np.array_equal(E["ok"][req_base:req_base + k], S["ok"]), g
        for t in range(3):
            assert np.array_equal(E["vac"][t][a:b], S["vac"][t]), (g, t)
        req_base += k
    assert req_base == len(E["t"])
    assert m_sum == m_all


def test_ensemble_step_matches_single_networks():
    sizes = [300, 1200, 50]
    _say("ensemble neuron step with per-network growth rates", sizes)
    off, boxes, xyz, sig, va, ve, vi = _ensemble(sizes, 7)
    n = int(off[-1])
    nu = np.array([1e-4, 3e-4, 5e-5])
    p = StepParams.izhikevich(input_mode=1)
    tgt = np.full(n, 0.7)
    big = BarnesHutHIP(boxes[:, :3].min(0), boxes[:, 3:].max(0), seed=9)
    big.init_networks(off, boxes, xyz, sig)
    big.neurons_init(p, tgt)
    big.neurons_set_network_growth(np.repeat(nu, 3))
    for s in range(20):
        big.neurons_step(s)
    E = big.neurons_get()
    big.close()
    for g in range(len(sizes)):
        a, b = int(off[g]), int(off[g + 1])
        one = BarnesHutHIP(boxes[g, :3], boxes[g, 3:], seed=network_seed(9, g))
        one.init(xyz[a:b], sig[a:b])
        pg = StepParams.izhikevich(input_mode=1)
        pg.nu[0] = pg.nu[1] = pg.nu[2] = nu[g]
        one.neurons_init(pg, tgt[a:b])
        for s in range(20):
            one.neurons_step(s)
        S = one.neurons_get()
        one.close()
        for key in ("x", "u", "ca", "fired"):
            assert np.array_equal(E[key][a:b], S[key]), (g, key)
        assert np.array_equal(E["delta"].reshape(3, n)[:, a:b], S["delta"].reshape(3, -1)), g


def test_ensemble_rejects_bad_layouts():
    off, boxes, xyz, sig, va, ve, vi = _ensemble([10, 20], 3)
    g = BarnesHutHIP(boxes[:, :3].min(0), boxes[:, 3:].max(0))
    with pytest.raises(BHError):
        g.init_networks(np.array([0, 10, 10, 30], np.uint64), np.vstack([boxes[:1], boxes]), xyz, sig)  # empty network
    with pytest.raises(BHError):
        g.init_networks(np.array([1, 10, 30], np.uint64), boxes, xyz, sig)  # offsets[0] != 0
    bad = boxes.copy()
    bad[1, 3] = bad[1, 0] - 1.0
    with pytest.raises(BHError):
        g.init_networks(off, bad, xyz, sig)  # min > max
    moved = xyz.copy()
    moved[15] = boxes[0, 3:] + 1.0  # a neuron of network 1 outside its box
    with pytest.raises(BHError):
        g.init_networks(off, boxes, moved, sig)
    g.close()


pytestmark = pytest.mark.gpu
