"""Statistics of the counter-based background input (rl_step_params.input_mode 1).

The device draws each neuron's Normal(mean, stddev) input of a step from its own Philox4x32-10 counter
(neuron id inside its network, step, tag 3) by a Box-Muller transform.  The reference draws this
input from Boost's normal distribution (NormalActivityInputCPU.cpp:35), whose stream is parity
unpinned, so this mode is checked for its distribution instead of against values.

The input is read back exactly through the step itself: with k1 = k2 = k3 = 0, a = b = 0, u = 0,
one substep and no spike threshold in reach, a step sets x to x + input.
"""
import sys

import numpy as np
import pytest
from scipy import stats

from relearn_amd import BarnesHutHIP, StepParams
from relearn_amd.workloads import parameter_sweep, uniform_cube

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _probe_params(mean, sd):
    p = StepParams.izhikevich(input_mode=1, input_mean=mean, input_stddev=sd)
    p.a = p.b = 0.0
    p.k1 = p.k2 = p.k3 = 0.0
    p.h = 1
    p.v_spike = 1e300
    return p


def _inputs(g, n, steps):
    """Inputs of `steps` consecutive steps (rows), recovered from the x differences."""
    out = np.zeros((steps, n))
    prev = np.zeros(n)
    for s in range(steps):
        g.neurons_step(s)
        x = g.neurons_get()["x"]
        out[s] = x - prev
        prev = x
    return out


def test_input_stream_is_normal_and_independent():
    n, steps, mean, sd = 1_000_000, 3, 5.0, 2.0
    _say("counter-based input statistics,", n, "neurons x", steps, "steps")
    net = uniform_cube(n, 10_000.0, 3)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=21)
    g.init(net.xyz, net.signal)
    g.neurons_init(_probe_params(mean, sd), np.full(n, 0.7), x=np.zeros(n), u=np.zeros(n))
    z = (_inputs(g, n, steps) - mean) / sd
    g.close()
    for s in range(steps):
        # mean and standard deviation within about 5 standard errors, and the whole distribution
        assert abs(z[s].mean()) < 5.0 / np.sqrt(n), s
        assert abs(z[s].std() - 1.0) < 5.0 / np.sqrt(2 * n), s
        ks = stats.kstest(z[s], "norm")
        _say("step", s, "mean", z[s].mean(), "std", z[s].std(), "KS p", ks.pvalue)
        assert ks.pvalue > 1e-3, s
        # no correlation between neighbouring neurons of a step
        assert abs(np.corrcoef(z[s][:-1], z[s][1:])[0, 1]) < 5.0 / np.sqrt(n), s
    # no correlation between steps of a neuron, and a new draw every step (the FP32 transform
    # quantises z to about 1e-7, so equal values of two steps are rare coincidences)
    assert abs(np.corrcoef(z[0], z[1])[0, 1]) < 5.0 / np.sqrt(n)
    assert abs(np.corrcoef(z[1], z[2])[0, 1]) < 5.0 / np.sqrt(n)
    assert np.mean(z[0] == z[1]) < 1e-3


def test_input_stream_follows_mean_and_stddev_and_network_seeds():
    # other parameters scale the same draws; each network of an ensemble draws its own stream
    ens = parameter_sweep(4, n_min=20_000, n_max=20_001, seed=2)
    n = ens.n
    g = BarnesHutHIP(ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0), seed=5)
    g.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
    g.neurons_init(_probe_params(-3.0, 0.5), np.full(n, 0.7), x=np.zeros(n), u=np.zeros(n))
    v = _inputs(g, n, 1)[0]
    g.close()
    assert abs(v.mean() + 3.0) < 5 * 0.5 / np.sqrt(n)
    assert abs(v.std() - 0.5) < 5 * 0.5 / np.sqrt(2 * n)
    off = ens.offsets.astype(np.int64)
    first = [v[off[k]:off[k] + 1000] for k in range(4)]
    for a in range(4):
        for b in range(a + 1, 4):
            assert not np.allclose(first[a], first[b])
            assert abs(np.corrcoef(first[a], first[b])[0, 1]) < 0.2
