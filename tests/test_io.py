"""The reference's neuron and network text formats (relearn_amd.io) on the committed fixtures."""
import os

import numpy as np

from relearn_amd import io

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_network_heads_parse():
    # the first data line of both files is "0 1\t0 2\t1\t1": target 1 <- source 2 (ids from 1)
    n = 51842
    off, src, w, pl = io.read_in_synapses(os.path.join(GOLDEN, "network_head_in.txt"), n)
    assert int(off[-1]) == 40 and int(off[1]) > 0 and src[0] == 1 and w[0] == 1 and pl[0] == 1
    oo, tgt, ow, opl = io.read_out_synapses(os.path.join(GOLDEN, "network_head_out.txt"), n)
    # both heads list the same 40 lines: the edge sets agree whichever way they are grouped
    t_in = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    s_out = np.repeat(np.arange(n), np.diff(oo.astype(np.int64)))
    assert sorted(zip(t_in.tolist(), src.tolist())) == sorted(zip(tgt.tolist(), s_out.tolist()))


def test_network_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n, m = 500, 4000
    tgt, src = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.choice([-2, -1, 1, 3], m)
    order = np.argsort(tgt, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(tgt, minlength=n))]).astype(np.uint64)
    io.write_in_synapses(tmp_path / "in.txt", off, src[order], w[order], step=7)
    o2, s2, w2, p2 = io.read_in_synapses(tmp_path / "in.txt", n)
    assert np.array_equal(o2, off) and np.array_equal(s2, src[order]) and np.array_equal(w2, w[order])
    assert np.all(p2 == 1)
    io.write_out_synapses(tmp_path / "out.txt", off, src[order], w[order])
    oo, tt, ww, _ = io.read_out_synapses(tmp_path / "out.txt", n)
    s_out = np.repeat(np.arange(n), np.diff(oo.astype(np.int64)))
    assert sorted(zip(tt.tolist(), s_out.tolist(), ww.tolist())) == sorted(zip(tgt.tolist(), src.tolist(), w.tolist()))


def test_neurons_round_trip(tmp_path):
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    xyz, sig = d["xyz"][:3000], d["signal"][:3000].copy()
    sig[::7] = 1
    io.write_neurons(tmp_path / "pos.txt", xyz, sig, box_min=d["box_min"], box_max=d["box_max"])
    r = io.read_neurons(tmp_path / "pos.txt")
    assert np.allclose(r["xyz"], xyz, atol=5e-7) and np.array_equal(r["signal"], sig)
    assert np.allclose(r["box_min"], d["box_min"]) and np.allclose(r["box_max"], d["box_max"])


def test_reference_positions_head_parses():
    # header and first 200 neurons of the reference's input/positions.txt: the same numbers as the
    # positions fixture, and the subdomain box of its header
    r = io.read_neurons(os.path.join(GOLDEN, "positions_head.txt"))
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    assert r["xyz"].shape == (200, 3)
    assert np.array_equal(r["xyz"], d["xyz"][:200]) and np.array_equal(r["signal"], d["signal"][:200])
    assert np.array_equal(r["box_min"], d["box_min"]) and np.array_equal(r["box_max"], d["box_max"])
    assert r["area"][0] == "lh_bankssts"
