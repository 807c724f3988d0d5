"""Level-1 candidate lists exported by the device (rl_bh_candidate_lists) against the oracle.

FP64: the same nodes in the same order and probabilities within 1e-13 relative (the device's FP64
exp and the C library's may differ in the last bits; the sums and picks built from them are checked
against the oracle in test_gpu_parity.py).  FP32-mixed: the same nodes (acceptance decisions inside
the guard band are re-made in FP64) and probabilities within 2e-5 relative, the tolerance that the
FP32 distance, FP32 exp2 and FP32 product allow.
"""
import sys

import numpy as np
import pytest

from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [RL_PREC_FP64, RL_PREC_FP32_MIXED])
def test_candidate_lists_match_oracle(oracle, precision):
    print("[gpu-case] candidate lists vs oracle, precision", precision, file=sys.stderr, flush=True)
    net = uniform_cube(30_000, 3000.0, 51)
    t = oracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh)
    g = BarnesHutHIP(net.box_min, net.box_max, precision=precision)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    sources = np.random.default_rng(5).choice(net.n, 64, replace=False).astype(np.uint32)
    lists = g.candidate_lists(sources, capacity=8192)
    worst = 0.0
    for s, (nodes, probs, count) in zip(sources, lists):
        on, op, total, fb, _ = t.candidate_list(0, net.xyz[s], int(s), int(net.signal[s]), 0.3, 750.0)
        assert count == len(on) == len(nodes)
        assert np.array_equal(nodes.astype(np.uint64), on)
        rel = np.abs(probs - op) / np.maximum(op, 1e-300)
        worst = max(worst, float(rel.max()) if len(rel) else 0.0)
    assert worst <= (1e-13 if precision == RL_PREC_FP64 else 2e-5), worst
    g.close()
    t.close()
