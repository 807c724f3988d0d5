"""Size-independent properties at the full precision workload: 16M neurons uniform in a 25,000 um
cube, 80/20 E/I, vacant axons Poisson(2), vacant dendrites Poisson(1.5) per type, theta 0.3,
sigma 750 um.  The oracle cannot run this size in seconds, so these checks stand in for it:

- octree: every neuron is exactly one leaf, every inner node's counts are the sums of its children's,
  the root holds all vacancies, leaves hold their neuron's position, every centroid lies in its
  node's box, and an inner node's centroid times its count is the sum of its children's (to the
  rounding of the FP64 sums);
- search: no source picks itself, every pick is a leaf whose dendrites of the source's type were
  vacant, searches stop at the first failure (BarnesHutBase.h:283-289), and FP32-mixed picks differ
  from FP64 picks in at most 1e-4 of the searches;
- commit: per (target, type) the accepted requests never exceed the vacancies and the vacancies drop
  by exactly the accepted count, per source by one axon per accepted request.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu

N, SIDE = 16_000_000, 25_000.0
NONE = 0xFFFFFFFF


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def net():
    return uniform_cube(N, SIDE, 0)


def _check_tree(tr, net):
    m = len(tr["level"])
    nch = tr["nchild"].astype(np.int64)
    link = tr["link"].astype(np.int64)
    leaf = nch == 0
    assert int(leaf.sum()) == N
    assert np.array_equal(np.sort(link[leaf]), np.arange(N))
    inner = np.nonzero(~leaf)[0]
    # children are contiguous: child k of node i is link[i] + k
    parent = np.repeat(inner, nch[inner])
    child = np.repeat(link[inner], nch[inner]) + (np.arange(len(parent)) - np.repeat(np.cumsum(nch[inner]) - nch[inner], nch[inner]))
    assert len(child) == m - 1 and np.array_equal(np.sort(child), np.arange(1, m))
    for key, vac in (("cnt_exc", net.vac_exc), ("cnt_inh", net.vac_inh)):
        cnt = tr[key].astype(np.int64)
        sums = np.bincount(parent, weights=cnt[child].astype(np.float64), minlength=m).astype(np.int64)
        assert np.array_equal(sums[inner], cnt[inner]), key
        assert int(cnt[0]) == int(vac.astype(np.int64).sum()), key
        assert np.array_equal(cnt[leaf], vac[link[leaf]].astype(np.int64)), key
    for key, ck in (("pos_exc", "cnt_exc"), ("pos_inh", "cnt_inh")):
        pos, cnt = tr[key], tr[ck].astype(np.float64)
        has = cnt > 0
        lo, hi = tr["box_lo"], tr["box_hi"]
        assert np.all((pos[has] >= lo[has]) & (pos[has] <= hi[has])), key
        # leaves hold their neuron's position; an inner node's centroid times its count is the sum of
        # its children's, up to the rounding of the FP64 sums
        lh = leaf & has
        assert np.array_equal(pos[lh], net.xyz[link[lh]]), key
        ih = inner[has[inner]]
        for d in range(3):
            w = pos[child, d] * cnt[child]
            acc = np.bincount(parent, weights=w, minlength=m)
            scale = np.bincount(parent, weights=np.abs(w), minlength=m)
            assert np.all(np.abs(pos[ih, d] * cnt[ih] - acc[ih]) <= 1e-12 * scale[ih] + 1e-300), (key, d)


def _check_search(g, net, slots):
    tgt, tests, cnt = g.get_search_slots(slots)
    va = net.vac_axon.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(va)])
    sn = np.repeat(np.arange(N), va)
    j = np.arange(slots) - off[sn]
    c = cnt.astype(np.int64)
    assert np.all(c <= va)
    found = j < c[sn]
    assert np.all(tgt[found] != NONE)
    # the first failure ends a neuron's searches: its later slots stay unsearched
    assert np.all(tgt[j > c[sn]] == NONE)
    picked = tgt[found].astype(np.int64)
    src = sn[found]
    assert np.all(picked != src)
    exc = net.signal[src] == 0
    assert np.all(net.vac_exc[picked[exc]] > 0) and np.all(net.vac_inh[picked[~exc]] > 0)
    return tgt


def _check_commit(g, net, st):
    m = st["requests"]
    t, s, sig, _ = g.get_requests(m)
    ok = g.get_commit_result(m).astype(bool)
    assert int(ok.sum()) == st["created"]
    va, ve, vi = g.get_vacancies()
    acc_exc = np.bincount(t[ok & (sig == 0)], minlength=N)
    acc_inh = np.bincount(t[ok & (sig != 0)], minlength=N)
    acc_ax = np.bincount(s[ok], minlength=N)
    assert np.all(acc_exc <= net.vac_exc) and np.all(acc_inh <= net.vac_inh)
    assert np.array_equal(ve.astype(np.int64), net.vac_exc.astype(np.int64) - acc_exc)
    assert np.array_equal(vi.astype(np.int64), net.vac_inh.astype(np.int64) - acc_inh)
    assert np.array_equal(va.astype(np.int64), net.vac_axon.astype(np.int64) - acc_ax)


def test_16m_update_properties(net):
    slots = int(net.vac_axon.sum())
    picks = {}
    for prec in (RL_PREC_FP64, RL_PREC_FP32_MIXED):
        _say("16M update properties, precision", prec)
        g = BarnesHutHIP(net.box_min, net.box_max, seed=1, precision=prec)
        g.init(net.xyz, net.signal)
        g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
        if prec == RL_PREC_FP64:
            _check_tree(g.export_tree(), net)
            _say("tree checked")
        nreq = g.find_target_neurons(0)
        picks[prec] = _check_search(g, net, slots)
        _say("search checked,", nreq, "requests")
        created = g.commit(0)
        _check_commit(g, net, {"requests": nreq, "created": created})
        _say("commit checked,", created, "created")
        g.close()
    a, b = picks[RL_PREC_FP64], picks[RL_PREC_FP32_MIXED]
    div = float(np.mean(a != b))
    _say("FP32-mixed vs FP64 pick divergence", div)
    assert div <= 1e-4
