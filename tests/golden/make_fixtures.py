"""Writes the committed fixtures of tests/golden/ (run in the build container, where the reference is mounted).

positions_51842.npz: the 51,842 real neuron positions of the reference's input/positions.txt (a data file
of the reference's own test inputs), parsed from its text columns <id> <x> <y> <z> <area> <type> into
float64 positions, a 0/1 signal type (ex -> 0, in -> 1) and the subdomain box stated in its header.
network_51842.npz: the in-network of the same neurons, input/rank_0_in_network.txt (425,706 plastic
synapses of weight 1 on rank 0), as target and source ids (from 0) of every line.
network_head_{in,out}.txt: the first 40 data lines of input/rank_0_in_network.txt and
input/rank_0_out_network.txt with their header lines, for the text-format readers.
positions_head.txt: the header and first 200 data lines of input/positions.txt, for the neuron-file
reader.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/relearn/input/positions.txt"


def positions_fixture(path=SRC):
    xyz, sig = [], []
    box_max = None
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                if "Local subdomain 0 boundaries" in line:
                    tail = line.split("-")[-1].strip().strip("()")
                    box_max = [float(v) for v in tail.split(",")]
                continue
            parts = line.split()
            if len(parts) < 6:
                continue
            xyz.append([float(parts[1]), float(parts[2]), float(parts[3])])
            sig.append(0 if parts[-1] == "ex" else 1)
    xyz = np.asarray(xyz, dtype=np.float64)
    return xyz, np.asarray(sig, dtype=np.uint8), np.asarray(box_max, dtype=np.float64)


def network_fixture(path="/root/reference/relearn/input/rank_0_in_network.txt"):
    tgt, src = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            p = line.split()
            if len(p) < 6:
                continue
            assert p[0] == "0" and p[2] == "0" and p[4] == "1" and p[5] == "1"
            tgt.append(int(p[1]) - 1)
            src.append(int(p[3]) - 1)
    return np.asarray(tgt, np.uint32), np.asarray(src, np.uint32)


def head(path, lines):
    out, data = [], 0
    with open(path) as fh:
        for line in fh:
            out.append(line)
            if not line.startswith("#"):
                data += 1
                if data == lines:
                    break
    return "".join(out)


def main():
    xyz, sig, box_max = positions_fixture()
    out = os.path.join(HERE, "positions_51842.npz")
    np.savez_compressed(out, xyz=xyz, signal=sig, box_min=np.zeros(3), box_max=box_max)
    print(out, xyz.shape, box_max, os.path.getsize(out))
    tgt, src = network_fixture()
    out = os.path.join(HERE, "network_51842.npz")
    np.savez_compressed(out, target=tgt, source=src)
    print(out, tgt.shape, os.path.getsize(out))
    for kind in ("in", "out"):
        out = os.path.join(HERE, f"network_head_{kind}.txt")
        with open(out, "w") as fh:
            fh.write(head(f"/root/reference/relearn/input/rank_0_{kind}_network.txt", 40))
        print(out, os.path.getsize(out))
    out = os.path.join(HERE, "positions_head.txt")
    with open(out, "w") as fh:
        fh.write(head(SRC, 200))
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    sys.exit(main())
