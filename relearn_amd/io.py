"""The reference's text formats for neurons and networks, read and written on the host.

Files written here are read by the reference's NeuronIO and files it writes are read here, so a
network grown on the device can be continued or analysed by the CPU code and vice versa.

  neurons (NeuronIO::read_neuron_positions_and_signals / write_neuron_positions_and_signals,
      source/io/NeuronIO.cpp:227-303 and 439-545):
      '#' header lines, then one line per neuron: <id> <x> <y> <z> <area> <type>, ids from 1, type
      'ex' or 'in'.
  in / out synapses (NeuronIO::read_in_synapses / write_in_synapses, NeuronIO.cpp:793-1100):
      '#' header lines, then one line per edge: <target rank> <target id> <source rank> <source id>
      <weight> <plastic>, ids from 1.  One rank here: every rank column is 0.

The device keeps the in-edges as a CSR by target (rl_net_init); in-edge files map onto it directly,
out-edge files are the same edges grouped by source.
"""
from __future__ import annotations

import re

import numpy as np


def read_neurons(path):
    """Positions, signal types (0 excitatory, 1 inhibitory), area names and the box of the header
    (the first local subdomain's boundaries, else the positions' bounding box)."""
    ids, xyz, sig, area = [], [], [], []
    box = None
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                if "Local subdomain 0 boundaries" in line:
                    groups = re.findall(r"\(([^)]*)\)", line)
                    if len(groups) >= 2:
                        box = (np.array([float(v) for v in groups[0].split(",")]),
                               np.array([float(v) for v in groups[1].split(",")]))
                continue
            parts = line.split()
            if len(parts) < 6:
                continue
            ids.append(int(parts[0]))
            xyz.append([float(parts[1]), float(parts[2]), float(parts[3])])
            area.append(parts[4])
            sig.append(0 if parts[-1] == "ex" else 1)
    ids = np.asarray(ids, dtype=np.int64)
    order = np.argsort(ids, kind="stable")
    if not np.array_equal(ids[order], np.arange(1, len(ids) + 1)):
        raise ValueError(f"{path}: neuron ids must be 1..n")
    xyz = np.asarray(xyz, dtype=np.float64)[order]
    sig = np.asarray(sig, dtype=np.uint8)[order]
    area = [area[i] for i in order]
    if box is None:
        box = (xyz.min(axis=0), xyz.max(axis=0))
    return {"xyz": xyz, "signal": sig, "area": area, "box_min": box[0], "box_max": box[1]}


def write_neurons(path, xyz, signal, area=None, box_min=None, box_max=None):
    xyz = np.asarray(xyz, dtype=np.float64)
    n = len(xyz)
    lo = xyz.min(axis=0) if box_min is None else np.asarray(box_min, dtype=np.float64)
    hi = xyz.max(axis=0) if box_max is None else np.asarray(box_max, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(f"# {n} of {n}\n")
        for d, c in enumerate("xyz"):
            fh.write(f"# Minimum {c}: {lo[d]:g}\n")
        for d, c in enumerate("xyz"):
            fh.write(f"# Maximum {c}: {hi[d]:g}\n")
        fh.write("# <local id> <pos x> <pos y> <pos z> <area> <type>\n")
        fh.write("# Number of local subdomains: 1\n")
        fh.write(f"# Local subdomain 0 boundaries ({lo[0]:g}, {lo[1]:g}, {lo[2]:g}) - ({hi[0]:g}, {hi[1]:g}, {hi[2]:g})\n")
        for i in range(n):
            a = "area" if area is None else area[i]
            t = "in" if signal[i] else "ex"
            fh.write(f"{i + 1}\t{xyz[i, 0]:.6f} {xyz[i, 1]:.6f} {xyz[i, 2]:.6f}\t{a}\t{t}\n")


def _read_edges(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) < 6:
                continue
            rows.append((int(parts[0]), int(parts[1]), int(parts[2]), int(parts[3]), int(parts[4]), int(parts[5])))
    a = np.asarray(rows, dtype=np.int64).reshape(-1, 6)
    if np.any(a[:, 0] != 0) or np.any(a[:, 2] != 0):
        raise ValueError(f"{path}: edges of other ranks are not supported (one rank)")
    if np.any(a[:, 4] == 0):
        raise ValueError(f"{path}: weight 0")
    return a


def _csr(key, other, w, plastic, n):
    if len(key) and (key.min() < 0 or key.max() >= n or other.min() < 0 or other.max() >= n):
        raise ValueError("neuron id outside 1..n")
    order = np.argsort(key, kind="stable")
    off = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=n))]).astype(np.uint64)
    return off, other[order].astype(np.uint32), w[order].astype(np.int32), plastic[order].astype(np.uint8)


def read_in_synapses(path, n):
    """In-edges as a CSR by target: (offsets n + 1, sources, weights, plastic flags), ids from 0."""
    a = _read_edges(path)
    return _csr(a[:, 1] - 1, a[:, 3] - 1, a[:, 4], a[:, 5], n)


def read_out_synapses(path, n):
    """Out-edges as a CSR by source: (offsets n + 1, targets, weights, plastic flags), ids from 0.
    (The file lists <target> <source> columns like the in-edge file.)"""
    a = _read_edges(path)
    return _csr(a[:, 3] - 1, a[:, 1] - 1, a[:, 4], a[:, 5], n)


def _write_edges(path, key, other, w, plastic, step, n, target_first):
    with open(path, "w") as fh:
        fh.write(f"# Total number neurons: {n}\n# Local number neurons: {n}\n# Number MPI ranks: 1\n")
        fh.write(f"# Current simulation step: {step}\n")
        fh.write("# <target rank> <target neuron id>\t<source rank> <source neuron id>\t<weight>\t<plastic>\n")
        for k, o, ww, p in zip(key.tolist(), other.tolist(), w.tolist(), plastic.tolist()):
            t, s = (k, o) if target_first else (o, k)
            fh.write(f"0 {t + 1}\t0 {s + 1}\t{ww}\t{p}\n")


def write_in_synapses(path, in_offsets, sources, weights, plastic=None, step=0):
    """The device's in-edge CSR (rl_net_get_edges) as the reference's in-network file."""
    off = np.asarray(in_offsets, dtype=np.int64)
    n = len(off) - 1
    t = np.repeat(np.arange(n), np.diff(off))
    pl = np.ones(len(t), np.uint8) if plastic is None else np.asarray(plastic)
    _write_edges(path, t, np.asarray(sources, dtype=np.int64), np.asarray(weights), pl, step, n, True)


def write_out_synapses(path, in_offsets, sources, weights, plastic=None, step=0):
    """The same edges grouped by source, as the reference's out-network file."""
    off = np.asarray(in_offsets, dtype=np.int64)
    n = len(off) - 1
    t = np.repeat(np.arange(n), np.diff(off))
    s = np.asarray(sources, dtype=np.int64)
    pl = np.ones(len(t), np.uint8) if plastic is None else np.asarray(plastic)
    order = np.lexsort((t, s))
    _write_edges(path, s[order], t[order], np.asarray(weights)[order], pl[order], step, n, False)
