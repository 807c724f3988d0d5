"""The reference's neuron and network text formats (relearn_amd.io) on the committed fixtures."""
import os

import numpy as np
import pytest

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
    # 17 significant digits: the written file reads back to the same doubles, box included (a box
    # rounded to 6 digits would put the largest x of the fixture, 141.0981, outside the box)
    assert np.array_equal(r["xyz"], xyz) and np.array_equal(r["signal"], sig)
    assert np.array_equal(r["box_min"], d["box_min"]) and np.array_equal(r["box_max"], d["box_max"])
    assert np.all(r["xyz"] <= r["box_max"]) and np.all(r["xyz"] >= r["box_min"])
    # random doubles, both layouts
    rng = np.random.default_rng(3)
    xyz2 = rng.uniform(0.0, 1000.0, (500, 3))
    sig2 = (rng.random(500) < 0.2).astype(np.uint8)
    for area in (None, ["a%d" % (i % 3) for i in range(500)]):
        io.write_neurons(tmp_path / "pos2.txt", xyz2, sig2, area=area)
        r2 = io.read_neurons(tmp_path / "pos2.txt")
        assert np.array_equal(r2["xyz"], xyz2) and np.array_equal(r2["signal"], sig2)
        assert np.array_equal(r2["box_min"], xyz2.min(axis=0)) and np.array_equal(r2["box_max"], xyz2.max(axis=0))
        if area is not None:
            assert r2["area"] == area


def test_reference_positions_head_parses():
    # header and first 200 neurons of the reference's input/positions.txt: the same numbers as the
    # positions fixture, and the subdomain box of its header
    r = io.read_neurons(os.path.join(GOLDEN, "positions_head.txt"))
    d = np.load(os.path.join(GOLDEN, "positions_51842.npz"))
    assert r["xyz"].shape == (200, 3)
    assert np.array_equal(r["xyz"], d["xyz"][:200]) and np.array_equal(r["signal"], d["signal"][:200])
    assert np.array_equal(r["box_min"], d["box_min"]) and np.array_equal(r["box_max"], d["box_max"])
    assert r["area"][0] == "lh_bankssts"


def test_calcium_file(tmp_path):
    # the layout of CalciumIO: '<rank>:<id> <initial> <target>' lines and one 'default' line; ids of
    # other ranks fall back to the default (test_calcium_io in the reference checks the same cases)
    p = tmp_path / "ca.txt"
    p.write_text("# comment\ndefault 0.1 0.7\n0:1 0.25 0.5\n3:2 9 9\n-1:3 0.125 0.875\n0:4;0:5 0.3 0.6\n"
                 "not a line\n")
    init, tgt = io.read_calcium(p, 6)
    assert np.array_equal(init, [0.25, 0.1, 0.125, 0.3, 0.3, 0.1])
    assert np.array_equal(tgt, [0.5, 0.7, 0.875, 0.6, 0.6, 0.7])
    # group names (areas of the older neuron layout)
    groups = io.groups_from_areas(["x", "y", "x", "y"])
    p.write_text("default 0 0.7\ny 0.5 0.2\n")
    init, tgt = io.read_calcium(p, 4, groups)
    assert np.array_equal(init, [0, 0.5, 0, 0.5]) and np.array_equal(tgt, [0.7, 0.2, 0.7, 0.2])
    # errors: an id twice, no default for an unnamed neuron, id 0
    for text in ("0:1 1 1\n0:1 2 2\n", "0:1 1 1\n", "default 1 1\n0:0 1 1\n"):
        p.write_text(text)
        with pytest.raises(ValueError):
            io.read_calcium(p, 2)
    # write -> read is exact
    rng = np.random.default_rng(0)
    a, b = rng.random(50), rng.random(50)
    io.write_calcium(p, a, b)
    a2, b2 = io.read_calcium(p, 50)
    assert np.array_equal(a, a2) and np.array_equal(b, b2)


def test_synaptic_elements_file(tmp_path):
    # 15 values per line in the order of SynapticElementsIO: per type (axons, excitatory, inhibitory
    # dendrites) min calcium, growth rate, retract ratio, min elements, max elements
    rng = np.random.default_rng(2)
    n = 25
    vals = rng.uniform(0.0, 10.0, (n, 15))
    default = rng.uniform(0.0, 10.0, 15)
    lines = ["default " + " ".join(f"{v:.6g}" for v in default)]
    for i in range(n):
        if i % 4 != 3:
            lines.append(f"0:{i + 1} " + " ".join(f"{v:.6g}" for v in vals[i]))
    p = tmp_path / "se.txt"
    p.write_text("\n".join(lines) + "\n")
    r = io.read_synaptic_elements(p, n)
    want = vals.copy()
    want[3::4] = default
    for k, f in enumerate(io.ELEMENT_FIELDS):
        assert r[f].shape == (3, n)
        for t in range(3):
            assert np.allclose(r[f][t], want[:, 5 * t + k], rtol=1e-5)
    # write -> read is exact
    io.write_synaptic_elements(p, r)
    r2 = io.read_synaptic_elements(p, n)
    for f in io.ELEMENT_FIELDS:
        assert np.array_equal(r[f], r2[f])


def test_neuron_groups_file(tmp_path):
    p = tmp_path / "groups.txt"
    p.write_text("# groups\n1 a b\nc 2 3\n3 a\n")
    of, names = io.read_neuron_groups(p, 3)
    assert names == ["default", "a", "b", "c"]
    assert of == [[0, 1, 2], [0, 3], [0, 3, 1]]
