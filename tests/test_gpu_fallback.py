"""The reference's probability fallbacks on the device, against the oracle.

Kernel::create_probability_interval (source/algorithm/Kernel/Kernel.h:90-134) replaces the Gaussian
probabilities by n / d when their sum underflows to 0 (Kernel.h:109-126).  Two workloads reach it:

- sparse: 2,000 neurons uniform in a 10^6 um cube with sigma = 750 um.  Neighbours are ~50,000 um
  (~67 sigma) apart, so every Gaussian underflows in FP64 (beyond ~27 sigma) and every list takes the
  n / d branch, at level 1 and in the descents.
- lattice: neurons on a jittered grid 15 sigma apart.  exp(-225) is a normal FP64 number, so the
  FP64 path keeps the Gaussian, but the FP32 kernel (exp2 of -325) underflows to 0: the FP32-mixed
  search must hand every source to the FP64 re-evaluation and end with the FP64 picks.

Both search kernels (cooperative, and the per-lane one of RL_FLAG_EXACT_ORDER) and both precisions
must give the oracle's picks except for draws within 1e-12 of a cumulative boundary.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from relearn_amd import RL_FLAG_EXACT_ORDER, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
from relearn_amd.workloads import Network, uniform_cube
from tests.test_gpu_parity import _check_search

pytestmark = pytest.mark.gpu

SIGMA = 750.0


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _lattice(per_side, spacing, seed):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(per_side)] * 3, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    xyz = (g + 0.5 + rng.uniform(-0.2, 0.2, g.shape)) * spacing
    n = len(xyz)
    sig = np.ones(n, np.uint8)
    sig[rng.permutation(n)[: math.ceil(0.8 * n)]] = 0
    side = per_side * spacing
    return Network(xyz, sig, rng.poisson(2.0, n).astype(np.uint32), rng.poisson(1.5, n).astype(np.uint32),
                   rng.poisson(1.5, n).astype(np.uint32), (0.0, 0.0, 0.0), (side, side, side))


def _run(oracle, net, precision, flags, seed, step):
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    ref = t.find_targets(net.signal, net.vac_axon, sigma=SIGMA, seed=seed, step=step, threads=0)
    t.close()
    g = BarnesHutHIP(net.box_min, net.box_max, sigma=SIGMA, seed=seed, precision=precision, flags=flags)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    nreq = g.find_target_neurons(step)
    st = g.last_search_stats()
    tgt, tests, cnt = g.get_search_slots(int(net.vac_axon.astype(np.uint64).sum()))
    g.close()
    return ref, nreq, st, tgt, tests, cnt


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_search_inverse_distance_fallback(oracle, precision, flags):
    _say("fallback n/d: sparse 2000 neurons in a 1e6 um cube, precision", precision, "flags", flags)
    net = uniform_cube(2000, 1.0e6, 41)
    ref, nreq, st, tgt, tests, cnt = _run(oracle, net, precision, flags, seed=43, step=2)
    # the oracle took the n / d branch (and searched a real workload: most searches find a target)
    assert ref["stats"]["fallbacks"] > 0
    assert ref["stats"]["found"] > 0.5 * ref["stats"]["searches"]
    # the device handed the lists whose Gaussian sum is 0 to the per-lane fallback pass
    assert st["lane_sources"] > 0
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("fallbacks", ref["stats"]["fallbacks"], "lane sources", st["lane_sources"], "near-boundary", near,
         "mismatches", len(bad), bad[:5])
    assert not bad
    assert nreq == int(cnt.sum())


@pytest.mark.parametrize("flags", [0, RL_FLAG_EXACT_ORDER])
@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_search_fp32_underflow_redone_in_fp64(oracle, precision, flags):
    _say("fp32 underflow: lattice 15 sigma apart, precision", precision, "flags", flags)
    net = _lattice(13, 15.0 * SIGMA, 47)
    ref, nreq, st, tgt, tests, cnt = _run(oracle, net, precision, flags, seed=53, step=1)
    # FP64 keeps the Gaussian everywhere
    assert ref["stats"]["fallbacks"] == 0
    if precision == RL_PREC_FP32_MIXED:
        # every FP32 level-1 sum is 0: the sources were redone in FP64
        assert st["lane_sources"] > 0.5 * np.count_nonzero(net.vac_axon)
    elif flags == 0:
        assert st["lane_sources"] == 0
    bad, near = _check_search(ref, tgt, tests, cnt, net)
    _say("lane sources", st["lane_sources"], "slow descents", st["slow_descents"], "near-boundary", near,
         "mismatches", len(bad), bad[:5])
    assert not bad
    assert nreq == int(cnt.sum())
