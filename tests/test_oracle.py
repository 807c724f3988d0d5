"""Pins the CPU oracle (oracle/bh_oracle.cpp) against known answers, closed forms and brute force.

Runs on the CPU only.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from tests.helpers import brute_octree, deep_cluster, subtree_leaves


# ------------------------------------------------------------------ random number streams

def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32-10
    assert [int(v) for v in oracle.philox([0, 0, 0, 0], [0, 0])] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert [int(v) for v in oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert [int(v) for v in oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                          [0xA4093822, 0x299F31D0])] == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_mt19937_known_answer(oracle):
    # C++ standard [rand.predef]: the 10000th invocation of a default-constructed mt19937 yields 4123659995
    assert oracle.mt19937_kth(5489, 10000) == 4123659995


def test_pick_draw_addresses_distinct(oracle):
    seen = set()
    for neuron in range(4):
        for j in range(3):
            for level in range(1, 4):
                seen.add(oracle.pick_draw(7, neuron, 5, j, level))
    assert len(seen) == 36
    # the step and the seed are part of the address
    assert oracle.pick_draw(7, 1, 5, 0, 1) != oracle.pick_draw(7, 1, 6, 0, 1)
    assert oracle.pick_draw(7, 1, 5, 0, 1) != oracle.pick_draw(8, 1, 5, 0, 1)


def test_counter_draws_are_uniform(oracle):
    # the pick draws and commit priorities of consecutive counters (neurons, axon ordinals, levels,
    # steps) are uniform 32-bit words: Kolmogorov-Smirnov on 2e5 draws, every bit balanced, no
    # correlation between neighbouring counters
    from scipy import stats
    draws = {
        "neuron": [oracle.pick_draw(7, i, 5, 0, 1) for i in range(100_000)],
        "axon": [oracle.pick_draw(7, 3, 5, j, 1) for j in range(50_000)] + [oracle.pick_draw(7, 4, 5, j, 2) for j in range(50_000)],
        "step": [oracle.pick_draw(9, 11, s, 0, 1) for s in range(100_000)],
        "priority": [oracle.commit_priority(7, i, 5, 0) >> 32 for i in range(100_000)],
    }
    for name, d in draws.items():
        u = np.asarray(d, dtype=np.float64) / 2.0**32
        assert stats.kstest(u, "uniform").pvalue > 1e-4, name
        bits = (np.asarray(d, dtype=np.uint64)[:, None] >> np.arange(32, dtype=np.uint64)) & np.uint64(1)
        assert np.all(np.abs(bits.mean(0) - 0.5) < 5 * 0.5 / np.sqrt(len(d))), name
        assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 5 / np.sqrt(len(d)), name


def test_draw_to_uniform_boost_mapping(oracle):
    # boost uniform_real_distribution on a 32-bit engine: k / 2^32 * next, always < next
    for nxt in (1.0, np.nextafter(3.7, 4.0), 1e-300, 12345.678):
        for k in (0, 1, 2**31, 2**32 - 1):
            r = oracle.draw_to_uniform(k, nxt)
            assert r == Fraction(k, 2**32) * Fraction(nxt) or abs(r - k / 2**32 * nxt) <= np.spacing(nxt)
            assert r < nxt
    assert oracle.draw_to_uniform(0, 5.0) == 0.0


# ------------------------------------------------------------------ picker

def test_picker_known_answers(oracle):
    # test/probability_picker/test_probability_picker.cpp:122-145
    assert oracle.pick_index([0.0, 2.0, 0.0, 1.421, 1.0], 1.0) == 1
    assert oracle.pick_index([0.0, 2.0, 0.0, 1.421, 1.0], 4.0) == 4
    assert oracle.pick_index([1.0, 0.0, 0.0, 1.0], 1.0) == 0
    assert oracle.pick_index([1.0, 0.0, 0.0, 1.0], np.nextafter(1.0, 2.0)) == 3


def test_picker_exceptions(oracle):
    # test_probability_picker.cpp:36-50: the reference throws; the oracle returns -1
    assert oracle.pick_index([], 1.0) == -1
    assert oracle.pick_index([1.0], -0.056) == -1
    assert oracle.pick_index([0.0], 1.0) == -1
    assert oracle.pick_index([0.0, 0.0, -1.0, 1.0], 1.0) == -1


def test_picker_lower_bound_equivalence(oracle):
    # test_probability_picker.cpp:72-115: index == lower_bound(prefix_sum, r) - 1, zero draw -> 0,
    # draws above the sum -> last index
    rng = np.random.default_rng(1)
    for _ in range(50):
        k = int(rng.integers(1, 100))
        p = rng.random(k) * 3.0
        p[0] += 0.001
        prefix = np.concatenate([[0.0], np.cumsum(p)[:-1]])
        total = float(np.sum(p))
        for _ in range(20):
            r = float(rng.uniform(np.nextafter(0.0, 1.0), total))
            idx = oracle.pick_index(p, r)
            # sequential partial sums, compared like the reference's loop
            s, ref = 0.0, None
            for i, v in enumerate(p):
                s += v
                if s >= r:
                    ref = i
                    break
            assert idx == ref
            assert idx == int(np.searchsorted(prefix, r, side="left")) - 1 or abs(prefix[idx] - r) < 1e-12 * total
        assert oracle.pick_index(p, 0.0) == 0
        assert oracle.pick_index(p, total * 1.5) == k - 1


def test_picker_zero_draw_skips_zero_probabilities(oracle):
    # ProbabilityPicker.h:56-70: a zero draw picks the first strictly positive entry
    assert oracle.pick_index([0.0, 0.0, 2.0, 1.0], 0.0) == 2


# ------------------------------------------------------------------ hypot

def _exact_norm(x, y, z):
    s = Fraction(x) ** 2 + Fraction(y) ** 2 + Fraction(z) ** 2
    return math.sqrt(s.numerator / s.denominator) if s < 1e300 else float("inf")


def test_hypot3_close_to_exact(oracle):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        v = rng.normal(size=3) * 10 ** rng.uniform(-3, 4)
        h = oracle.hypot3(*v)
        exact = _exact_norm(*v)
        assert abs(h - exact) <= 4 * np.spacing(exact)
    assert oracle.hypot3(3.0, 4.0, 0.0) == 5.0
    assert oracle.hypot3(0.0, 0.0, 0.0) == 0.0
    assert oracle.hypot3(-2.5, 0.0, 0.0) == 2.5


# ------------------------------------------------------------------ octree

def _small_net(n, side, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.random((n, 3)) * side
    return xyz


@pytest.mark.parametrize("n,seed", [(2, 0), (3, 1), (17, 2), (300, 3)])
def test_tree_matches_brute_force(oracle, n, seed):
    xyz = _small_net(n, 1000.0, seed)
    t = oracle.Tree(xyz, [0, 0, 0], [1000, 1000, 1000])
    e = t.export()
    ref = brute_octree(xyz, [0, 0, 0], [1000, 1000, 1000])
    assert len(ref) == t.num_nodes
    for i, (level, lo, hi, leaf, octant, nc) in enumerate(ref):
        assert e["level"][i] == level
        assert np.array_equal(e["box_lo"][i], lo) and np.array_equal(e["box_hi"][i], hi)
        assert e["nchild"][i] == nc
        assert e["octant"][i] == octant
        if nc == 0:
            assert e["link"][i] == leaf
        assert e["side"][i] == max(hi - lo)


def test_tree_single_neuron_is_root_leaf(oracle):
    t = oracle.Tree(np.array([[1.0, 2.0, 3.0]]), [0, 0, 0], [10, 10, 10])
    e = t.export()
    assert t.num_nodes == 1 and e["nchild"][0] == 0 and e["link"][0] == 0


def test_tree_insertion_order_independent(oracle):
    xyz = _small_net(500, 1000.0, 5)
    perm = np.random.default_rng(6).permutation(500)
    a = oracle.Tree(xyz, [0, 0, 0], [1000] * 3).export()
    b = oracle.Tree(xyz[perm], [0, 0, 0], [1000] * 3).export()
    assert np.array_equal(a["level"], b["level"]) and np.array_equal(a["box_lo"], b["box_lo"])
    leaves = a["nchild"] == 0
    assert np.array_equal(a["link"][leaves], perm[b["link"][leaves]])


def test_tree_rejects_duplicates_and_outside(oracle):
    with pytest.raises(oracle.OracleError):
        oracle.Tree(np.array([[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]]), [0, 0, 0], [10, 10, 10])
    with pytest.raises(oracle.OracleError):
        oracle.Tree(np.array([[11.0, 1.0, 1.0]]), [0, 0, 0], [10, 10, 10])


def test_upward_pass_counts_and_centroids(oracle):
    rng = np.random.default_rng(7)
    n = 2000
    xyz = rng.random((n, 3)) * 1000
    ve = rng.poisson(1.5, n).astype(np.uint32)
    vi = rng.poisson(1.5, n).astype(np.uint32)
    en = (rng.random(n) > 0.1).astype(np.uint8)
    t = oracle.Tree(xyz, [0, 0, 0], [1000] * 3)
    t.update(ve, vi, en, threads=4)
    e = t.export()
    leaves = subtree_leaves(e)
    we = np.where(en == 1, ve, 0).astype(np.float64)
    wi = np.where(en == 1, vi, 0).astype(np.float64)
    for i in range(t.num_nodes):
        ids = np.array(leaves[i])
        for cnt_key, pos_key, has_key, w in (("cnt_exc", "pos_exc", "has_exc", we), ("cnt_inh", "pos_inh", "has_inh", wi)):
            total = w[ids].sum()
            assert e[cnt_key][i] == total
            if e["nchild"][i] == 0:
                assert e[has_key][i] == 1 and np.array_equal(e[pos_key][i], xyz[ids[0]])
                continue
            assert e[has_key][i] == (1 if total > 0 else 0)
            if total > 0:
                cen = (xyz[ids] * w[ids][:, None]).sum(0) / total
                assert np.allclose(e[pos_key][i], cen, rtol=1e-12, atol=1e-9)
                assert np.all(e[pos_key][i] >= e["box_lo"][i]) and np.all(e[pos_key][i] <= e["box_hi"][i])


def test_upward_pass_child_order_exact(oracle):
    # the centroid is formed from the children in octant order: acc += pos_c * (double)n_c
    rng = np.random.default_rng(8)
    n = 400
    xyz = rng.random((n, 3)) * 1000
    ve = rng.poisson(1.5, n).astype(np.uint32)
    vi = rng.poisson(1.5, n).astype(np.uint32)
    t = oracle.Tree(xyz, [0, 0, 0], [1000] * 3)
    t.update(ve, vi)
    e = t.export()
    for i in np.nonzero(e["nchild"] > 0)[0]:
        f, nc = int(e["link"][i]), int(e["nchild"][i])
        acc = np.zeros(3)
        tot = 0
        for c in range(f, f + nc):
            if e["cnt_exc"][c] > 0:
                acc = acc + e["pos_exc"][c] * float(e["cnt_exc"][c])
            tot += int(e["cnt_exc"][c])
        if tot:
            q = np.minimum(np.maximum(acc / float(tot), e["box_lo"][i]), e["box_hi"][i])
            assert np.array_equal(q, e["pos_exc"][i])


def test_real_positions_fixture_tree(oracle):
    import os

    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "positions_51842.npz"))
    t = oracle.Tree(d["xyz"], d["box_min"], d["box_max"])
    e = t.export()
    leaves = e["nchild"] == 0
    assert leaves.sum() == len(d["xyz"])
    assert sorted(e["link"][leaves].tolist()) == list(range(len(d["xyz"])))
    # every inner node holds >= 2 neurons, every leaf lies in its box
    lv = subtree_leaves(e)
    assert all(len(lv[i]) >= 2 for i in np.nonzero(~leaves)[0])
    p = d["xyz"][e["link"][leaves]]
    assert np.all(p >= e["box_lo"][leaves]) and np.all(p <= e["box_hi"][leaves])


# ------------------------------------------------------------------ search

def test_candidate_list_small_theta_is_all_leaves(oracle):
    # with theta -> 0 no inner node passes the criterion, so the list is every leaf with vacant
    # dendrites except the source's own, in descending Morton order (reverse BFS leaf order)
    rng = np.random.default_rng(9)
    n = 200
    xyz = rng.random((n, 3)) * 1000
    ve = rng.poisson(1.5, n).astype(np.uint32)
    t = oracle.Tree(xyz, [0, 0, 0], [1000] * 3)
    t.update(ve, ve)
    e = t.export()
    src = 17
    nodes, prob, total, fb, tests = t.candidate_list(0, xyz[src], src, 0, 1e-9, 750.0)
    # leaves in Morton order = DFS order with octants ascending
    order = []

    def dfs(i):
        nc = int(e["nchild"][i])
        if nc == 0:
            order.append(i)
            return
        f = int(e["link"][i])
        for c in range(f, f + nc):
            dfs(c)

    dfs(0)
    expect = [i for i in reversed(order) if e["cnt_exc"][i] > 0 and e["link"][i] != src]
    assert list(nodes) == expect
    d = np.linalg.norm(xyz[e["link"][nodes]] - xyz[src], axis=1)
    assert np.allclose(prob, np.exp(-(d / 750.0) ** 2) * e["cnt_exc"][nodes], rtol=1e-13)
    assert total == pytest.approx(prob.sum(), rel=1e-14)


def test_candidate_list_theta_accepts_far_cells(oracle):
    # every accepted inner node satisfies side < theta * d, every expanded one does not; the list
    # covers each vacant leaf (other than the source) exactly once
    rng = np.random.default_rng(10)
    n = 3000
    xyz = rng.random((n, 3)) * 1000
    ve = rng.poisson(1.5, n).astype(np.uint32)
    t = oracle.Tree(xyz, [0, 0, 0], [1000] * 3)
    t.update(ve, ve)
    e = t.export()
    leaves = subtree_leaves(e)
    src = 5
    nodes, prob, total, fb, tests = t.candidate_list(0, xyz[src], src, 0, 0.3, 750.0)
    covered = []
    for i in nodes:
        d = np.linalg.norm(e["pos_exc"][i] - xyz[src])
        if e["nchild"][i] > 0:
            assert e["side"][i] < 0.3 * d * (1 + 1e-12)
        covered.extend(leaves[i])
    vacant = {i for i in range(n) if ve[i] > 0 and i != src}
    covered_vacant = [i for i in covered if ve[i] > 0]
    assert sorted(covered_vacant) == sorted(vacant)
    assert src not in covered
    assert len(nodes) < len(vacant)


def test_find_targets_statistics(oracle):
    # picks follow p ~ exp(-d^2/sigma^2) * n: nearer vacant neurons are chosen more often
    rng = np.random.default_rng(11)
    n = 400
    xyz = rng.random((n, 3)) * 1000
    ve = np.ones(n, np.uint32)
    sig = np.zeros(n, np.uint8)
    va = np.zeros(n, np.uint32)
    va[0] = 3000
    t = oracle.Tree(xyz, [0, 0, 0], [1000] * 3)
    t.update(ve, ve)
    r = t.find_targets(sig, va, seed=3, threads=1, sigma=150.0)
    tg = r["target"][: r["count"][0]]
    assert r["count"][0] == 3000 and np.all(tg != 0)
    d = np.linalg.norm(xyz[tg] - xyz[0], axis=1)
    dall = np.linalg.norm(xyz[1:] - xyz[0], axis=1)
    w = np.exp(-(dall / 150.0) ** 2)
    assert abs(np.mean(d) - np.sum(dall * w) / np.sum(w)) < 0.1 * np.sum(dall * w) / np.sum(w)


def _fallback_tree(oracle, counts):
    # source 0 at (1000, 1000, 1000); A, B, C at distances 3000, 4000 and 5000 um along the axes
    xyz = np.array([[1000.0, 1000.0, 1000.0], [4000.0, 1000.0, 1000.0], [1000.0, 5000.0, 1000.0],
                    [1000.0, 1000.0, 6000.0]])
    ve = np.asarray(counts, np.uint32)
    t = oracle.Tree(xyz, [0, 0, 0], [10000.0] * 3)
    t.update(ve, ve)
    return t, xyz


def test_candidate_list_inverse_distance_fallback(oracle):
    # Kernel::create_probability_interval (Kernel.h:109-126): when every Gaussian underflows
    # (sigma = 1 um against 3000-5000 um) the probabilities become n / d.  Hand-computed: A (n = 2,
    # d = 3000) 2/3000, B (n = 1, d = 4000) 1/4000, C (n = 3, d = 5000) 3/5000.
    t, xyz = _fallback_tree(oracle, [5, 2, 1, 3])
    e = t.export()
    nodes, prob, total, fb, tests = t.candidate_list(0, xyz[0], 0, 0, 1e-6, 1.0)
    assert fb
    got = {int(e["link"][i]): p for i, p in zip(nodes, prob)}
    assert set(got) == {1, 2, 3}
    assert got[1] == pytest.approx(2 / 3000, rel=1e-15)
    assert got[2] == pytest.approx(1 / 4000, rel=1e-15)
    assert got[3] == pytest.approx(3 / 5000, rel=1e-15)
    assert total == pytest.approx(2 / 3000 + 1 / 4000 + 3 / 5000, rel=1e-15)
    # a Gaussian that does not underflow keeps the Gaussian (sigma = 3000: A = 2 e^-1)
    nodes, prob, total, fb, _ = t.candidate_list(0, xyz[0], 0, 0, 1e-6, 3000.0)
    assert not fb
    got = {int(e["link"][i]): p for i, p in zip(nodes, prob)}
    assert got[1] == pytest.approx(2 * math.exp(-1.0), rel=1e-15)
    # only the source's own leaf has dendrites: the list is empty whatever the kernel
    t2, _ = _fallback_tree(oracle, [5, 0, 0, 0])
    nodes, prob, total, fb, _ = t2.candidate_list(0, xyz[0], 0, 0, 1e-6, 1.0)
    assert len(nodes) == 0 and total == 0.0


def test_find_targets_inverse_distance_picks(oracle):
    # With the n / d fallback the picks of one source split as n_A / d_A : n_B / d_B.  A (n = 1,
    # d = 1000) and B (n = 3, d = 3000) are equally likely; a 1/d^2 weight would give 3 : 1 and a
    # weight without n 3 : 1 the other way.
    xyz = np.array([[1000.0, 1000.0, 1000.0], [2000.0, 1000.0, 1000.0], [1000.0, 4000.0, 1000.0]])
    ve = np.array([0, 1, 3], np.uint32)
    t = oracle.Tree(xyz, [0, 0, 0], [10000.0] * 3)
    t.update(ve, ve)
    va = np.array([4000, 0, 0], np.uint32)
    r = t.find_targets(np.zeros(3, np.uint8), va, seed=5, threads=1, theta=1e-6, sigma=1.0)
    assert r["count"][0] == 4000 and r["stats"]["fallbacks"] > 0
    frac_a = float(np.mean(r["target"] == 1))
    assert abs(frac_a - 0.5) < 0.035   # 4.4 standard errors at 4000 draws
    # sigma large enough for the Gaussian: no fallback, A wins exp(-(1/3)^2) : 3 exp(-1)
    r = t.find_targets(np.zeros(3, np.uint8), va, seed=5, threads=1, theta=1e-6, sigma=3000.0)
    assert r["stats"]["fallbacks"] == 0
    pa, pb = math.exp(-1.0 / 9.0), 3.0 * math.exp(-1.0)
    assert abs(float(np.mean(r["target"] == 1)) - pa / (pa + pb)) < 0.035


def test_find_targets_deterministic_across_threads(oracle):
    rng = np.random.default_rng(12)
    n = 3000
    xyz = rng.random((n, 3)) * 2000
    va = rng.poisson(2, n).astype(np.uint32)
    ve = rng.poisson(1.5, n).astype(np.uint32)
    vi = rng.poisson(1.5, n).astype(np.uint32)
    sig = (rng.random(n) < 0.2).astype(np.uint8)
    t = oracle.Tree(xyz, [0, 0, 0], [2000] * 3)
    t.update(ve, vi)
    a = t.find_targets(sig, va, seed=4, threads=1)
    b = t.find_targets(sig, va, seed=4, threads=8)
    assert np.array_equal(a["target"], b["target"]) and np.array_equal(a["count"], b["count"])
    assert np.array_equal(a["node_tests"], b["node_tests"])
    c = t.find_targets(sig, va, seed=5, threads=8)
    assert not np.array_equal(a["target"], c["target"])
    # no autapse, no target without vacant dendrites of the source's type
    for i in range(n):
        o = int(a["offsets"][i])
        for tg in a["target"][o:o + a["count"][i]]:
            assert tg != i
            assert (ve if sig[i] == 0 else vi)[tg] > 0


# ------------------------------------------------------------------ commit

def _serial_commit(target, source, signal, order, ve, vi, va):
    ok = np.zeros(len(target), np.uint8)
    for i in order:
        vac = ve if signal[i] == 0 else vi
        if target[i] == source[i] or vac[target[i]] == 0:
            continue
        vac[target[i]] -= 1
        ok[i] = 1
    for i in range(len(target)):
        if ok[i]:
            va[source[i]] -= 1
    return ok


def test_commit_matches_serial_loop(oracle):
    rng = np.random.default_rng(13)
    n, m = 50, 400
    tgt = rng.integers(0, n, m).astype(np.uint32)
    src = rng.integers(0, n, m).astype(np.uint32)
    sig = rng.integers(0, 2, m).astype(np.uint8)
    jj = np.zeros(m, np.uint32)
    va = np.bincount(src, minlength=n).astype(np.uint32)
    ve = rng.poisson(2, n).astype(np.uint32)
    vi = rng.poisson(2, n).astype(np.uint32)
    order = rng.permutation(m)
    e1, i1, a1 = ve.copy(), vi.copy(), va.copy()
    ok, created = oracle.commit(tgt, src, sig, jj, e1, i1, a1, order=order)
    e2, i2, a2 = ve.copy(), vi.copy(), va.copy()
    ok2 = _serial_commit(tgt, src, sig, order, e2, i2, a2)
    assert np.array_equal(ok, ok2) and created == ok2.sum()
    assert np.array_equal(e1, e2) and np.array_equal(i1, i2) and np.array_equal(a1, a2)
    # closed form: per (target, type) the accepted count is min(vacant, non-self requests)
    for t in range(n):
        for s, vac in ((0, ve), (1, vi)):
            sel = (tgt == t) & (sig == s)
            nonself = int(np.sum(sel & (src != t)))
            assert int(ok[sel].sum()) == min(int(vac[t]), nonself)


def test_reference_shuffle_is_permutation(oracle):
    p = oracle.reference_shuffle(1000, 12345)
    assert sorted(p.tolist()) == list(range(1000))
    assert not np.array_equal(p, np.arange(1000))


# ------------------------------------------------------------------ neuron step

def _izh_params(oracle, h=10):
    p = oracle.StepParams()
    p.a, p.b, p.c, p.d = 0.02, 0.2, -65.0, 8.0
    p.v_spike, p.k1, p.k2, p.k3 = 30.0, 0.04, 5.0, 140.0
    p.h = h
    p.tau_ca, p.beta_ca = 10000.0, 0.001
    p.eta[0], p.eta[1], p.eta[2] = 0.4, 0.1, 0.0
    p.nu[0], p.nu[1], p.nu[2] = 1e-4, 1e-4, 1e-4
    return p


def test_neuron_step_single_substep_closed_form(oracle):
    p = _izh_params(oracle, h=1)
    x = np.array([-65.0, -70.0, 25.0]); u = np.array([-13.0, -14.0, 2.0]); ca = np.array([0.5, 0.0, 0.7])
    inp = np.array([5.0, 3.0, 10.0]); fired = np.zeros(3, np.uint8); tgt = np.full(3, 0.7); delta = np.zeros(9)
    x0, u0, c0 = x.copy(), u.copy(), ca.copy()
    oracle.neuron_step(p, x, u, ca, inp, fired, tgt, delta)
    for i in range(3):
        xi = ((((0.04 * x0[i]) * x0[i]) + (5.0 * x0[i])) + 140.0 - u0[i]) + inp[i]
        ui = ((0.02 * 0.2) * x0[i]) - (0.02 * u0[i])
        xn, un = x0[i] + xi * 1.0, u0[i] + ui * 1.0
        if xn >= 30.0:
            assert fired[i] == 1 and x[i] == -65.0 and u[i] == un + 8.0
        else:
            assert fired[i] == 0 and x[i] == xn and u[i] == un
        cn = c0[i] + 1.0 * (c0[i] * (-1.0 / 10000.0) + 0.001 * fired[i])
        assert ca[i] == cn
    # growth: dz = 2 exp(-((ca - xi)/zeta)^2) - 1 with xi = (eta+eps)/2, zeta = (eta-eps)/1.6651...
    for t, eta in enumerate((0.4, 0.1, 0.0)):
        for i in range(3):
            eps = max(0.7, eta)
            dz = 2 * math.exp(-(((ca[i] - (eta + eps) * 0.5) / ((eta - eps) / 1.6651092223153955)) ** 2)) - 1
            assert delta[t * 3 + i] == pytest.approx(dz * 1e-4, rel=1e-12)


def test_growth_curve_zero_at_eta_and_target(oracle):
    p = _izh_params(oracle, h=1)
    p.tau_ca = 1e300  # keep calcium fixed
    for ca0, t in ((0.4, 0), (0.7, 0), (0.1, 1), (0.7, 1)):
        x = np.array([-65.0]); u = np.array([-13.0]); ca = np.array([ca0]); inp = np.array([0.0])
        fired = np.zeros(1, np.uint8); delta = np.zeros(3)
        oracle.neuron_step(p, x, u, ca, inp, fired, np.array([0.7]), delta)
        assert abs(delta[t]) < 1e-15


def test_elements_update_cases(oracle):
    # ElementBaseBase::update: growth keeps connections; shrinking below the connected count deletes
    grown = np.array([3.5, 3.5, 3.5, 0.5]); delta = np.array([1.0, -0.25, -1.8, -2.0])
    conn = np.array([2, 2, 2, 0], np.uint32); vac = np.zeros(4, np.uint32)
    deleted = oracle.elements_update(grown, delta, conn, vac)
    new_count = 3.5 + -1.8
    assert grown.tolist() == [4.5, 3.25, (1 - 0.0) * (new_count - math.floor(new_count)) + math.floor(new_count), 0.0]
    assert vac.tolist() == [2, 1, 0, 0]
    assert conn.tolist() == [2, 2, 1, 0]
    assert deleted.tolist() == [0, 0, 1, 0]
    assert np.all(delta == 0)


def test_synaptic_input_matches_dense_product(oracle):
    # closed form: input = background + conductance * A f, where A[t, s] sums the weights of the
    # edges s -> t (repeated pairs add up) and f is the fired vector; the source's status counts,
    # never the target's
    rng = np.random.default_rng(21)
    n, m = 60, 700
    tgt = rng.integers(0, n, m)
    src = rng.integers(0, n, m)
    w = rng.integers(-3, 4, m).astype(np.int32)
    order = np.argsort(tgt, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(tgt, minlength=n))]).astype(np.uint64)
    fired = (rng.random(n) < 0.4).astype(np.uint8)
    bg = rng.normal(5.0, 2.0, n)
    A = np.zeros((n, n))
    np.add.at(A, (tgt, src), w)
    got = oracle.synaptic_input(off, src[order], w[order], fired, 0.25, bg)
    want = bg + 0.25 * (A @ fired.astype(np.float64))
    assert np.array_equal(got, want)
    # without background input the result is the synaptic term alone; no edges give 0
    assert np.array_equal(oracle.synaptic_input(off, src[order], w[order], fired, 0.25), 0.25 * (A @ fired))
    empty = np.zeros(n + 1, np.uint64)
    assert np.array_equal(oracle.synaptic_input(empty, np.zeros(0, np.uint32), np.zeros(0, np.int32), fired, 2.0),
                          np.zeros(n))


def _consistent_network(oracle, n, m, seed):
    """A random pair-merged graph with connected counts that match it (axons: out units; excitatory /
    inhibitory dendrites: in units of positive / negative weight), signal = sign of the source."""
    rng = np.random.default_rng(seed)
    signal = (rng.random(n) < 0.25).astype(np.uint8)
    tgt = rng.integers(0, n, m)
    src = rng.integers(0, n, m)
    w = np.where(signal[src] != 0, -1, 1).astype(np.int32)
    order = np.argsort(tgt, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(tgt, minlength=n))]).astype(np.uint64)
    g = oracle.Graph(n, off, src[order].astype(np.uint32), w[order])
    conn = np.zeros((3, n), np.uint32)
    eo, es, ew = g.export()
    t_of = np.repeat(np.arange(n), np.diff(eo.astype(np.int64)))
    np.add.at(conn[0], es, np.abs(ew))
    np.add.at(conn[1], t_of[ew > 0], ew[ew > 0])
    np.add.at(conn[2], t_of[ew < 0], -ew[ew < 0])
    return g, signal, conn, (off, src[order].astype(np.uint32), w[order])


def _pairs(graph):
    off, src, w = graph.export()
    t = np.repeat(np.arange(len(off) - 1), np.diff(off.astype(np.int64)))
    return {(int(a), int(b)): int(c) for a, b, c in zip(t, src, w)}


@pytest.mark.parametrize("etype", [0, 1, 2])
def test_deletion_conserves_counts(oracle, etype):
    # every deleted unit lowers one pair of the deleting neuron by one and disconnects exactly one
    # element of the partner; neurons delete exactly what they were asked to; others are untouched
    n = 300
    g, signal, conn, _ = _consistent_network(oracle, n, 3000, 1)
    before = _pairs(g)
    rng = np.random.default_rng(2)
    units = conn[etype].astype(np.int64)
    to_del = np.minimum(units, rng.integers(0, 4, n)).astype(np.uint32)
    c, v = conn.copy(), np.zeros((3, n), np.uint32)
    tot = g.delete(etype, to_del, signal, c, v, step=3, seed=9)
    assert tot == int(to_del.sum())
    after = _pairs(g)
    lost = {k: before[k] - after.get(k, 0) for k in before}
    per_neuron = np.zeros(n, np.int64)
    per_partner = np.zeros(n, np.int64)
    for (t, s), d in lost.items():
        if d == 0:
            continue
        assert abs(before[(t, s)]) >= abs(after.get((t, s), 0))
        mine, partner = (s, t) if etype == 0 else (t, s)
        if etype == 1:
            assert before[(t, s)] > 0
        if etype == 2:
            assert before[(t, s)] < 0
        per_neuron[mine] += abs(d)
        per_partner[partner] += abs(d)
    assert np.array_equal(per_neuron, to_del.astype(np.int64))
    # partner side: axons disconnect the target's dendrite of the source's sign, dendrites the source's axon
    if etype == 0:
        drop = np.zeros((3, n), np.int64)
        for (t, s), d in lost.items():
            drop[1 if signal[s] == 0 else 2][t] += abs(d)
        assert np.array_equal(conn.astype(np.int64) - c, drop) and np.array_equal(v.astype(np.int64), drop)
    else:
        assert np.array_equal(conn[0].astype(np.int64) - c[0], per_partner)
        assert np.array_equal(v[0].astype(np.int64), per_partner)
    g.close()


def test_deletion_all_or_nothing(oracle):
    n = 120
    g, signal, conn, _ = _consistent_network(oracle, n, 900, 3)
    before = _pairs(g)
    c, v = conn.copy(), np.zeros((3, n), np.uint32)
    assert g.delete(0, np.zeros(n, np.uint32), signal, c, v, step=0) == 0
    assert _pairs(g) == before
    # deleting every axon unit removes every pair
    assert g.delete(0, conn[0], signal, c, v, step=0) == int(conn[0].sum())
    assert _pairs(g) == {}
    assert np.all(c[1] == 0) and np.all(c[2] == 0)
    g.close()


def test_deletion_is_order_independent_and_uniform(oracle):
    # the drawn subset depends on the pairs, not on the order they were stored in; over many update
    # indices every unit of a neuron is drawn with frequency k/m
    n = 40
    _, signal, conn, (off, src, w) = _consistent_network(oracle, n, 400, 4)
    perm_src, perm_w = src.copy(), w.copy()
    for t in range(n):
        a, b = int(off[t]), int(off[t + 1])
        p = np.random.default_rng(t).permutation(b - a)
        perm_src[a:b], perm_w[a:b] = src[a:b][p], w[a:b][p]
    i = int(np.argmax(conn[0]))
    k, m = 2, int(conn[0][i])
    to_del = np.zeros(n, np.uint32)
    to_del[i] = k
    hits = {}
    trials = 3000
    for step in range(trials):
        res = []
        for s_, w_ in ((src, w), (perm_src, perm_w)):
            g = oracle.Graph(n, off, s_, w_)
            before = _pairs(g)
            g.delete(0, to_del, signal, conn.copy(), np.zeros((3, n), np.uint32), step=step, seed=11)
            after = _pairs(g)
            res.append({key: before[key] - after.get(key, 0) for key in before if before[key] != after.get(key, 0)})
            g.close()
        assert res[0] == res[1]
        for key, d in res[0].items():
            hits[key] = hits.get(key, 0) + abs(d)
    units = {key: abs(val) for key, val in _pairs(oracle.Graph(n, off, src, w)).items() if key[1] == i}
    assert sum(units.values()) == m
    for key, u in units.items():
        expect = trials * k * u / m
        assert abs(hits.get(key, 0) - expect) < 5 * math.sqrt(expect) + 5, (key, hits.get(key, 0), expect)


def test_deep_cluster_lists_reach_past_the_staging_ring(oracle):
    # the deep-tree workload of the GPU walk tests: level-1 lists of the cluster's sources hold nodes 27
    # or more levels below the root, deeper than the group walk's 24 LDS blocks of staged children
    # (the children of a node at stack level l, level l + 1 below the root, are staged into block l mod 24)
    net = deep_cluster(19)
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    lev = t.export()["level"]
    depth = t.depth
    deepest = 0
    for i in range(3000, len(net.xyz), 5):
        nodes, prob, total, fb, tests = t.candidate_list(0, net.xyz[i], i, int(net.signal[i]), 0.3, 1.0e12)
        if len(nodes):
            deepest = max(deepest, int(lev[nodes.astype(np.int64)].max()))
    t.close()
    assert depth >= 28
    assert deepest >= 27
