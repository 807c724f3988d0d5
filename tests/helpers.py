"""Shared test helpers: reference-style brute force on small inputs."""
from __future__ import annotations

import numpy as np


def brute_octree(xyz, box_min, box_max):
    """Capacity-1 PR octree by recursive midpoint partition (independent of insertion).

    Returns a list of nodes in breadth-first order (level, then octant order below each parent):
    (level, box_lo, box_hi, leaf_neuron or -1, octant, children count).
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    root = (0, np.array(box_min, float), np.array(box_max, float), np.arange(len(xyz)), 0)
    out = []
    frontier = [root]
    while frontier:
        nxt = []
        for level, lo, hi, ids, octant in frontier:
            if len(ids) <= 1:
                out.append((level, lo, hi, int(ids[0]) if len(ids) == 1 else -1, octant, 0))
                continue
            mid = (lo + hi) / 2.0
            pts = xyz[ids]
            code = ((pts[:, 0] >= mid[0]).astype(int) | ((pts[:, 1] >= mid[1]).astype(int) << 1)
                    | ((pts[:, 2] >= mid[2]).astype(int) << 2))
            kids = []
            for o in range(8):
                sel = ids[code == o]
                if len(sel) == 0:
                    continue
                clo, chi = lo.copy(), hi.copy()
                for d in range(3):
                    if (o >> d) & 1:
                        clo[d] = mid[d]
                    else:
                        chi[d] = mid[d]
                kids.append((level + 1, clo, chi, sel, o))
            out.append((level, lo, hi, -1, octant, len(kids)))
            nxt.extend(kids)
        frontier = nxt
    return out


def subtree_leaves(export):
    """For every node of an export (BFS, contiguous children), the list of leaf neuron ids below it."""
    m = len(export["level"])
    leaves = [None] * m
    for i in range(m - 1, -1, -1):
        nc = int(export["nchild"][i])
        if nc == 0:
            leaves[i] = [int(export["link"][i])]
        else:
            f = int(export["link"][i])
            acc = []
            for c in range(f, f + nc):
                acc.extend(leaves[c])
            leaves[i] = acc
    return leaves
