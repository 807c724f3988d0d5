"""The C++ host adapter (relearn_amd/host/barnes_hut_hip.hpp): the reference's Algorithm interface
(init, prepare_update_connectivity, update_connectivity) on top of the C ABI.

The CPU test checks that the example program links against the in-tree library and reports usage
errors without touching a GPU; the GPU test runs two plasticity updates through the adapter and
through the ctypes binding on the same inputs and requires identical synapses and vacancies.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from relearn_amd import RL_PREC_FP32_MIXED, BarnesHutHIP, build
from relearn_amd.workloads import uniform_cube


def _say(*a):
    print("[host-adapter]", *a, file=sys.stderr, flush=True)


def test_example_links_and_checks_usage():
    exe = build.HOST_EXAMPLE
    if not os.path.exists(exe):
        pytest.fail("update_example is not built (run __graft_entry__.build())")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "librelearn_bh.so" in out and "not found" not in out
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _write_input(path, net, steps, theta, sigma, seed, precision):
    with open(path, "wb") as f:
        f.write(np.array([net.n], np.uint64).tobytes())
        f.write(np.array([steps], np.uint32).tobytes())
        f.write(np.array([theta, sigma], np.float64).tobytes())
        f.write(np.array([seed], np.uint64).tobytes())
        f.write(np.array([precision], np.int32).tobytes())
        f.write(np.asarray(net.box_min, np.float64).tobytes())
        f.write(np.asarray(net.box_max, np.float64).tobytes())
        f.write(np.ascontiguousarray(net.xyz, np.float64).tobytes())
        f.write(np.ascontiguousarray(net.signal, np.uint8).tobytes())
        for v in (net.vac_axon, net.vac_exc, net.vac_inh):
            f.write(np.ascontiguousarray(v, np.uint32).tobytes())


def _read_output(path, n, steps):
    raw = open(path, "rb").read()
    off = 0
    per_step = []
    for _ in range(steps):
        cnt = int(np.frombuffer(raw, np.uint64, 1, off)[0])
        off += 8
        rec = np.frombuffer(raw, np.int32, 3 * cnt, off).reshape(-1, 3)
        off += 12 * cnt
        per_step.append(rec)
    vac = np.frombuffer(raw, np.uint32, 3 * n, off).reshape(3, n)
    return per_step, vac


@pytest.mark.gpu
def test_adapter_matches_binding(tmp_path):
    _say("adapter vs ctypes binding, 20k neurons, 2 updates")
    net = uniform_cube(20_000, 10_000.0 * (0.02 ** (1 / 3)), 31)
    steps, theta, sigma, seed = 2, 0.3, 750.0, 5
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(inp, net, steps, theta, sigma, seed, RL_PREC_FP32_MIXED)
    r = subprocess.run([build.HOST_EXAMPLE, inp, outp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    syn_adapter, vac_adapter = _read_output(outp, net.n, steps)

    g = BarnesHutHIP(net.box_min, net.box_max, theta=theta, sigma=sigma, seed=seed, precision=RL_PREC_FP32_MIXED)
    g.init(net.xyz, net.signal)
    va, ve, vi = net.vac_axon.copy(), net.vac_exc.copy(), net.vac_inh.copy()
    for step in range(steps):
        g.prepare_update_connectivity(va, ve, vi)
        nreq = g.find_target_neurons(step)
        created = g.commit(step)
        t, s, sig, _ = g.get_requests(nreq)
        ok = g.get_commit_result(nreq).astype(bool)
        ref = np.stack([t[ok].astype(np.int32), s[ok].astype(np.int32), np.where(sig[ok] == 1, -1, 1)], axis=1)
        assert created == ref.shape[0] > 0
        assert np.array_equal(syn_adapter[step], ref)
        va, ve, vi = g.get_vacancies()
    g.close()
    assert np.array_equal(vac_adapter, np.stack([va, ve, vi]))
