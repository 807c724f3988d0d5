"""The reference's text formats for neurons, networks, calcium and synaptic elements, read and written
on the host.

Files written here are read by the reference's readers and files it writes are read here, so a
network grown on the device can be continued or analysed by the CPU code and vice versa.

  neurons (NeuronIO::read_neuron_positions_and_signals / write_neuron_positions_and_signals_componentwise,
      source/io/NeuronIO.cpp:227-303 and 639-708):
      '#' header lines, then one line per neuron: <id> <x> <y> <z> <type>, ids from 1, type 'ex' or
      'in'.  The older layout with an area column, <id> <x> <y> <z> <area> <type> (the reference's
      input/positions.txt), is read too.  Values are written with 17 significant digits
      (max_digits10), so a written file reads back to the same doubles.
  neuron groups (NeuronIO::read_neuron_groups, NeuronIO.cpp:373-428): lines '<id> <group> ...' or
      '<group> <id> ...'.
  in / out synapses (NeuronIO::read_in_synapses / write_in_synapses, NeuronIO.cpp:793-1100):
      '#' header lines, then one line per edge: <target rank> <target id> <source rank> <source id>
      <weight> <plastic>, ids from 1.  One rank here: every rank column is 0.
  calcium (CalciumIO::load_initial_and_target_function, source/io/CalciumIO.h:37-49, CalciumIO.cpp:36-124):
      '<description> <initial calcium> <target calcium>' per line, or once 'default <initial> <target>'.
  synaptic elements (SynapticElementsIO::load_function_from_file, source/io/SynapticElementsIO.h:60,
      SynapticElementsIO.cpp:58-145): '<description>' followed by min calcium, growth rate, vacant retract
      ratio, min elements, max elements for axons, then excitatory, then inhibitory dendrites (15
      values), or once 'default' with the 15 values.
  A description (MonitorParser::parse_my_ids, NeuronIdParser::parse_multiple_description) is a ';'
  separated list of '<rank>:<id>' (ids from 1, rank -1 = this rank) and group names; entries of other
  ranks are ignored.  One rank here: this rank is 0.

The device keeps the in-edges as a CSR by target (rl_net_init); in-edge files map onto it directly,
out-edge files are the same edges grouped by source.
"""
from __future__ import annotations

import re

import numpy as np


def _g17(v):
    return format(float(v), ".17g")


def read_neurons(path):
    """Positions, signal types (0 excitatory, 1 inhibitory), area names (the older layout's column,
    else 'default') and the box of the header (the first local subdomain's boundaries, else the
    positions' bounding box)."""
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
            if len(parts) < 5:
                continue
            ids.append(int(parts[0]))
            xyz.append([float(parts[1]), float(parts[2]), float(parts[3])])
            area.append(parts[4] if len(parts) >= 6 else "default")
            sig.append(1 if parts[-1] == "in" else 0)
    ids = np.asarray(ids, dtype=np.int64)
    order = np.argsort(ids, kind="stable")
    if not np.array_equal(ids[order], np.arange(1, len(ids) + 1)):
        raise ValueError(f"{path}: neuron ids must be 1..n")
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)[order]
    sig = np.asarray(sig, dtype=np.uint8)[order]
    area = [area[i] for i in order]
    if box is None:
        box = (xyz.min(axis=0), xyz.max(axis=0))
    return {"xyz": xyz, "signal": sig, "area": area, "box_min": box[0], "box_max": box[1]}


def write_neurons(path, xyz, signal, area=None, box_min=None, box_max=None):
    """Writes the neuron file; with `area` in the older layout with an area column."""
    xyz = np.asarray(xyz, dtype=np.float64)
    n = len(xyz)
    lo = xyz.min(axis=0) if box_min is None else np.asarray(box_min, dtype=np.float64)
    hi = xyz.max(axis=0) if box_max is None else np.asarray(box_max, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(f"# {n} of {n}\n")
        for d, c in enumerate("xyz"):
            fh.write(f"# Minimum {c}: {_g17(lo[d])}\n")
        for d, c in enumerate("xyz"):
            fh.write(f"# Maximum {c}: {_g17(hi[d])}\n")
        fh.write("# <local id> <pos x> <pos y> <pos z> <area> <type>\n" if area is not None
                 else "# <local id> <pos x> <pos y> <pos z> <type>\n")
        fh.write("# Number of local subdomains: 1\n")
        fh.write(f"# Local subdomain 0 boundaries ({_g17(lo[0])}, {_g17(lo[1])}, {_g17(lo[2])}) - "
                 f"({_g17(hi[0])}, {_g17(hi[1])}, {_g17(hi[2])})\n")
        for i in range(n):
            t = "in" if signal[i] else "ex"
            a = "" if area is None else f" {area[i]}"
            fh.write(f"{i + 1} {_g17(xyz[i, 0])} {_g17(xyz[i, 1])} {_g17(xyz[i, 2])}{a} {t}\n")


def read_neuron_groups(path, n):
    """Groups of the neurons: (group ids per neuron, group names).  Every neuron is in group 0,
    'default', and in the groups its lines name."""
    names = ["default"]
    of = [[0] for _ in range(n)]

    def gid(name):
        if name not in names:
            names.append(name)
        return names.index(name)

    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0].isdigit():
                i = int(parts[0]) - 1
                if not 0 <= i < n:
                    raise ValueError(f"{path}: neuron id {parts[0]} outside 1..{n}")
                of[i].extend(gid(g) for g in parts[1:])
            else:
                g = gid(parts[0])
                for t in parts[1:]:
                    if not t.isdigit() or not 0 < int(t) <= n:
                        raise ValueError(f"{path}: {t} is not a neuron id in 1..{n}")
                    of[int(t) - 1].append(g)
    return of, names


def groups_from_areas(area):
    """Group name -> neuron ids (from 0) for the area column of the older neuron layout."""
    out = {}
    for i, a in enumerate(area):
        out.setdefault(a, []).append(i)
    return out


def _parse_ids(description, n, groups, rank=0):
    """Neuron ids (from 0) of this rank named by a description: ';'-separated '<rank>:<id>' entries
    (ids from 1, rank -1 = this rank) and group names."""
    ids = set()
    for piece in description.split(";"):
        if ":" in piece:
            r, _, i = piece.partition(":")
            try:
                r, i = int(r), int(i)
            except ValueError:
                continue
            if r == -1:
                r = rank
            if r < 0:
                continue
            if i <= 0:
                raise ValueError(f"neuron id 0 in {description!r}: ids start at 1")
            if r == rank:
                if i > n:
                    raise ValueError(f"neuron id {i} in {description!r} outside 1..{n}")
                ids.add(i - 1)
        elif groups is not None and piece in groups:
            ids.update(groups[piece])
    return sorted(ids)


def _read_per_neuron(path, n, width, groups):
    """Rows of `width` values per neuron from a description file with an optional default line."""
    default = None
    rows = {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("#"):
                continue
            parts = line.split()
            if len(parts) < 1 + width:
                continue
            try:
                vals = [float(v) for v in parts[1:1 + width]]
            except ValueError:
                continue
            if parts[0] == "default":
                if default is not None:
                    raise ValueError(f"{path}: more than one default line")
                default = vals
                continue
            for i in _parse_ids(parts[0], n, groups):
                if i in rows:
                    raise ValueError(f"{path}: neuron id {i + 1} assigned twice")
                rows[i] = vals
    out = np.empty((n, width))
    for i in range(n):
        v = rows.get(i, default)
        if v is None:
            raise ValueError(f"{path}: no value for neuron id {i + 1} and no default line")
        out[i] = v
    return out


def read_calcium(path, n, groups=None):
    """Initial and target calcium per neuron (n each) from a calcium file (--file-calcium).  groups:
    group name -> neuron ids, for descriptions that name groups."""
    v = _read_per_neuron(path, n, 2, groups)
    return v[:, 0].copy(), v[:, 1].copy()


def write_calcium(path, initial, target):
    with open(path, "w") as fh:
        fh.write("# <rank>:<neuron id> <initial calcium> <target calcium>\n")
        for i, (a, b) in enumerate(zip(np.asarray(initial).tolist(), np.asarray(target).tolist())):
            fh.write(f"0:{i + 1} {_g17(a)} {_g17(b)}\n")


ELEMENT_FIELDS = ("min_calcium", "nu", "retract_ratio", "min_elements", "max_elements")


def read_synaptic_elements(path, n, groups=None):
    """Per-neuron synaptic element parameters (--synaptic-elements-file): a dict of (3, n) arrays, type
    order axons, excitatory, inhibitory dendrites, for min_calcium, nu, retract_ratio, min_elements and
    max_elements."""
    v = _read_per_neuron(path, n, 15, groups).reshape(n, 3, 5)
    return {f: np.ascontiguousarray(v[:, :, k].T) for k, f in enumerate(ELEMENT_FIELDS)}


def write_synaptic_elements(path, params):
    """Writes a dict like read_synaptic_elements returns."""
    arrs = [np.asarray(params[f], dtype=np.float64) for f in ELEMENT_FIELDS]
    n = arrs[0].shape[1]
    with open(path, "w") as fh:
        fh.write("# <rank>:<neuron id> then per type (axons, excitatory, inhibitory dendrites): "
                 "min_calcium nu retract_ratio min_elements max_elements\n")
        for i in range(n):
            vals = [_g17(arrs[k][t, i]) for t in range(3) for k in range(5)]
            fh.write(f"0:{i + 1} " + " ".join(vals) + "\n")


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
