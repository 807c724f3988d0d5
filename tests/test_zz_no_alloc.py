"""The allocation-free commit (rl_bh_commit_requests_noalloc) against rl_bh_commit on the same requests.

Two calls in a row share the library's static scratch array, the second on a smaller network, so a
result that depended on the first call's leftovers would differ from the context's commit.
"""
import numpy as np
import pytest

from relearn_amd import RL_PREC_FP64, BarnesHutHIP
from relearn_amd._lib import BHError, commit_requests_noalloc, network_seed
from relearn_amd.workloads import uniform_cube

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime the library links (no second runtime in the process)."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def dev(self, a):
        a = np.ascontiguousarray(a)
        p = self.C.c_void_p()
        assert self.h.hipMalloc(self.C.byref(p), max(a.nbytes, 1)) == 0
        self.bufs.append((p.value, a.dtype, a.shape))
        assert self.h.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def host(self, ptr):
        for q, dt, shape in self.bufs:
            if q == ptr:
                out = np.empty(shape, dt)
                assert self.h.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
                return out
        raise KeyError(ptr)

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for q, _, _ in self.bufs:
            self.h.hipFree(q)
        self.bufs = []


def _case(n, side, seed, step, wait=True):
    print(f"no-alloc commit: {n} neurons, seed {seed}, step {step}, wait {wait}", flush=True)
    net = uniform_cube(n, side, seed=seed)
    g = BarnesHutHIP(net.box_min, net.box_max, seed=seed + 1, precision=RL_PREC_FP64)
    g.init(net.xyz, net.signal)
    g.prepare_update_connectivity(net.vac_axon, net.vac_exc, net.vac_inh)
    m = g.find_target_neurons(step)
    tgt, src, sig, j = g.get_requests(m)
    created_ref = g.commit(step)
    ok_ref = g.get_commit_result(m)
    va_ref, ve_ref, vi_ref = g.get_vacancies()
    g.close()
    assert m > 1000 and 0 < created_ref < m

    hip = _Hip()
    d_req = [hip.dev(tgt), hip.dev(src), hip.dev(sig), hip.dev(j)]
    d_vac = [hip.dev(net.vac_axon), hip.dev(net.vac_exc), hip.dev(net.vac_inh)]
    d_conn = [hip.dev(np.zeros(n, np.uint32)) for _ in range(3)]
    d_ok = hip.dev(np.full(m, 7, np.uint8))
    hip.sync()
    created = commit_requests_noalloc(n, m, *d_req, network_seed(seed + 1, 0), step, d_vac, d_conn, d_ok, wait=wait)
    hip.sync()
    ok = hip.host(d_ok)
    va, ve, vi = (hip.host(p) for p in d_vac)
    ca, ce, ci = (hip.host(p) for p in d_conn)
    hip.free()
    assert np.array_equal(ok, ok_ref)
    assert np.array_equal(va, va_ref) and np.array_equal(ve, ve_ref) and np.array_equal(vi, vi_ref)
    assert np.array_equal(ca, net.vac_axon - va_ref)
    assert np.array_equal(ce, net.vac_exc - ve_ref) and np.array_equal(ci, net.vac_inh - vi_ref)
    if wait:
        assert created == created_ref
    assert int(ok.sum()) == created_ref
    return m


def test_noalloc_commit_matches_context_commit_twice():
    m1 = _case(60_000, 3000.0, 11, 3)
    m2 = _case(8_000, 1500.0, 12, 5)
    assert m2 < m1
    _case(2_000, 1000.0, 13, 7, wait=False)


def test_noalloc_commit_rejects_sizes_beyond_the_scratch_array():
    with pytest.raises(BHError) as e:
        commit_requests_noalloc(1 << 19, 10, 0, 0, 0, 0, 0, 0, [0, 0, 0], None, 0)
    assert e.value.code == -7 and "262144" in str(e.value)
    with pytest.raises(BHError) as e:
        commit_requests_noalloc(1000, (1 << 20) + 1, 0, 0, 0, 0, 0, 0, [0, 0, 0], None, 0)
    assert e.value.code == -7
