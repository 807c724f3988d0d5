"""Seeded synthetic inputs for the Barnes-Hut connectivity update.

uniform_cube(): neurons uniform in [0, side)^3, the first ceil(0.8 n) of a random permutation
excitatory (BoxBasedRandomSubdomainAssignment places 80/20 E/I, source/sim/
BoxBasedRandomSubdomainAssignment.cpp:71-72), vacant axons ~ Poisson(2), vacant excitatory and
inhibitory dendrites ~ Poisson(1.5) each.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Network:
    xyz: np.ndarray        # (n, 3) float64
    signal: np.ndarray     # (n,) uint8, 0 excitatory, 1 inhibitory
    vac_axon: np.ndarray   # (n,) uint32
    vac_exc: np.ndarray
    vac_inh: np.ndarray
    box_min: tuple
    box_max: tuple

    @property
    def n(self):
        return self.xyz.shape[0]


def uniform_cube(n: int, side: float, seed: int = 0, axon_mean: float = 2.0, dend_mean: float = 1.5,
                 exc_fraction: float = 0.8) -> Network:
    rng = np.random.Generator(np.random.PCG64(seed))
    xyz = rng.random((n, 3)) * side
    n_exc = int(math.ceil(exc_fraction * n))
    perm = rng.permutation(n)
    signal = np.ones(n, np.uint8)
    signal[perm[:n_exc]] = 0
    vac_axon = rng.poisson(axon_mean, n).astype(np.uint32)
    vac_exc = rng.poisson(dend_mean, n).astype(np.uint32)
    vac_inh = rng.poisson(dend_mean, n).astype(np.uint32)
    return Network(xyz, signal, vac_axon, vac_exc, vac_inh, (0.0, 0.0, 0.0), (side, side, side))


# The configurations of the connectivity-update workloads: (neurons, cube side in micrometres)
CONFIGS = {
    "cfg0_1k": (1_000, 1_000.0),
    "cfg1_1m": (1_000_000, 10_000.0),
    "cfg2_16m": (16_000_000, 25_000.0),
    "cfg4_4m": (4_000_000, 16_000.0),
}
