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


@dataclass
class Ensemble:
    """Independent networks laid out one after another: network g owns neurons offsets[g]:offsets[g+1]
    and the box boxes[g] = (min xyz, max xyz); growth rates nu[g]."""
    offsets: np.ndarray
    boxes: np.ndarray
    xyz: np.ndarray
    signal: np.ndarray
    vac_axon: np.ndarray
    vac_exc: np.ndarray
    vac_inh: np.ndarray
    nu: np.ndarray

    @property
    def n(self) -> int:
        return int(self.offsets[-1])

    @property
    def num_networks(self) -> int:
        return len(self.offsets) - 1


def parameter_sweep(num_networks: int = 256, n_min: int = 2_000, n_max: int = 200_000, seed: int = 0,
                    nu_min: float = 5e-5, nu_max: float = 5e-4) -> Ensemble:
    """Ensemble of a parameter sweep: sizes log-uniform in [n_min, n_max], each network uniform in a
    cube of side 1000 * (n / 1000)^(1/3) micrometres, 80/20 E/I, vacancies Poisson(2) / (1.5) / (1.5),
    growth rates log-uniform in [nu_min, nu_max]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), num_networks)).astype(np.int64)
    sizes = np.maximum(sizes, 1)
    nets = [uniform_cube(int(k), 1000.0 * (k / 1000.0) ** (1.0 / 3.0), seed=seed * 100003 + g + 1)
            for g, k in enumerate(sizes)]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    boxes = np.array([list(nt.box_min) + list(nt.box_max) for nt in nets], np.float64)
    nu = np.exp(rng.uniform(math.log(nu_min), math.log(nu_max), num_networks))
    return Ensemble(offsets, boxes, np.concatenate([nt.xyz for nt in nets]), np.concatenate([nt.signal for nt in nets]),
                    np.concatenate([nt.vac_axon for nt in nets]), np.concatenate([nt.vac_exc for nt in nets]),
                    np.concatenate([nt.vac_inh for nt in nets]), nu)


# The uniform-cube workloads: (neurons, cube side in micrometres)
CONFIGS = {
    "cube_1k": (1_000, 1_000.0),
    "cube_1m": (1_000_000, 10_000.0),
    "cube_16m": (16_000_000, 25_000.0),
    "cube_4m": (4_000_000, 16_000.0),
}
