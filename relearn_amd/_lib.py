"""ctypes binding of librelearn_bh.so (the C ABI declared in include/relearn_bh.h).

The product path is the HIP library; there is no CPU fallback.  Loading fails loudly when the
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librelearn_bh.so")

RL_OK = 0
RL_PREC_FP64 = 0
RL_PREC_FP32_MIXED = 1
RL_FLAG_EXACT_ORDER = 1
RL_FLAG_WAVE_PER_SOURCE = 2
RL_FLAG_COUNT_WALK = 4
RL_FLAG_WALK_ALWAYS = 8

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)

EXPORTED = [
    "rl_bh_create", "rl_bh_destroy", "rl_bh_last_error", "rl_bh_init", "rl_bh_init_networks", "rl_bh_network_seed",
    "rl_bh_set_enabled",
    "rl_bh_prepare_update_connectivity", "rl_bh_prepare_resident", "rl_bh_prepare_update_connectivity_device",
    "rl_bh_find_target_neurons",
    "rl_bh_get_requests", "rl_bh_get_search_slots", "rl_bh_commit", "rl_bh_get_commit_result",
    "rl_bh_update_connectivity", "rl_bh_get_vacancies", "rl_bh_tree_info", "rl_bh_export_tree",
    "rl_bh_last_search_ms", "rl_bh_last_search_stats", "rl_neurons_init", "rl_neurons_step", "rl_neurons_commit_elements",
    "rl_neurons_get", "rl_neurons_set_network_growth", "rl_bh_graph_create", "rl_bh_graph_launch",
    "rl_bh_graph_destroy", "rl_bh_set_tuning", "rl_bh_get_tuning", "rl_bh_autotune", "rl_net_init",
    "rl_net_get_edges", "rl_bh_candidate_lists", "rl_neurons_commit_elements_at", "rl_neurons_set_element_params",
    "rl_bh_create_neurons", "rl_bh_commit_requests_noalloc",
]


class Config(C.Structure):
    _fields_ = [
        ("box_min", C.c_double * 3),
        ("box_max", C.c_double * 3),
        ("theta", C.c_double),
        ("sigma", C.c_double),
        ("mu", C.c_double),
        ("seed", C.c_uint64),
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("flags", C.c_int32),
        ("reserved", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("searches", C.c_uint64),
        ("requests", C.c_uint64),
        ("created", C.c_uint64),
        ("node_tests", C.c_uint64),
        ("node_tests_ref", C.c_uint64),
        ("ms_prepare", C.c_double),
        ("ms_search", C.c_double),
        ("ms_commit", C.c_double),
        ("ms_search_kernel", C.c_double),
        ("ms_total", C.c_double),
        ("deleted", C.c_uint64),
        ("net_capacity", C.c_uint64),
        ("net_regrown", C.c_uint64),
        ("lane_sources", C.c_uint64),
        ("slow_descents", C.c_uint64),
        ("walk_steps", C.c_uint64),
        ("walk_slots", C.c_uint64),
        ("walk_table_bytes", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("search_waves_per_cu", C.c_uint32),
        ("search_chunk", C.c_uint32),
        ("walk_min_sources", C.c_uint64),
        ("reserved", C.c_uint32 * 4),
    ]


class StepParams(C.Structure):
    _fields_ = [
        ("a", C.c_double),
        ("b", C.c_double),
        ("c", C.c_double),
        ("d", C.c_double),
        ("v_spike", C.c_double),
        ("k1", C.c_double),
        ("k2", C.c_double),
        ("k3", C.c_double),
        ("h", C.c_uint32),
        ("input_mode", C.c_uint32),
        ("input_mean", C.c_double),
        ("input_stddev", C.c_double),
        ("tau_ca", C.c_double),
        ("beta_ca", C.c_double),
        ("eta", C.c_double * 3),
        ("nu", C.c_double * 3),
        ("synapse_conductance", C.c_double),
        ("synaptic_input", C.c_uint32),
        ("reserved", C.c_uint32),
        ("retract_ratio", C.c_double * 3),
    ]

    @classmethod
    def izhikevich(cls, input_mode=0, input_mean=5.0, input_stddev=2.0, nu=1e-4):
        """The reference's defaults: Izhikevich a=0.02, b=0.2, c=-65, d=8, V_spike=30, k1=0.04, k2=5,
        k3=140, h=10 substeps; calcium tau 1e4, beta 1e-3; eta 0.4/0.1/0.0; growth rate nu."""
        p = cls()
        p.a, p.b, p.c, p.d = 0.02, 0.2, -65.0, 8.0
        p.v_spike, p.k1, p.k2, p.k3 = 30.0, 0.04, 5.0, 140.0
        p.h = 10
        p.input_mode = input_mode
        p.input_mean, p.input_stddev = input_mean, input_stddev
        p.tau_ca, p.beta_ca = 10000.0, 0.001
        p.eta[0], p.eta[1], p.eta[2] = 0.4, 0.1, 0.0
        p.nu[0] = p.nu[1] = p.nu[2] = nu
        return p


def _ptr(a, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


_lib = None


def load():
    """Loads librelearn_bh.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with relearn_amd/build.py (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.rl_bh_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.rl_bh_destroy.argtypes = [vp]
    L.rl_bh_last_error.argtypes = [vp]
    L.rl_bh_last_error.restype = C.c_char_p
    L.rl_bh_init.argtypes = [vp, C.c_uint64, _f64p, _u8p]
    L.rl_bh_init_networks.argtypes = [vp, C.c_uint32, _u64p, _f64p, _f64p, _u8p]
    L.rl_bh_network_seed.argtypes = [C.c_uint64, C.c_uint32]
    L.rl_bh_create_neurons.argtypes = [vp, C.c_uint64, _f64p, _u8p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.rl_neurons_set_network_growth.argtypes = [vp, _f64p]
    L.rl_neurons_set_element_params.argtypes = [vp, _f64p, _f64p]
    L.rl_bh_set_enabled.argtypes = [vp, _u8p]
    L.rl_bh_prepare_update_connectivity.argtypes = [vp, _u32p, _u32p, _u32p]
    L.rl_bh_prepare_resident.argtypes = [vp]
    L.rl_bh_prepare_update_connectivity_device.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
    L.rl_bh_find_target_neurons.argtypes = [vp, C.c_uint32, _u64p]
    L.rl_bh_get_requests.argtypes = [vp, C.c_uint64, _u32p, _u32p, _u8p, _u32p]
    L.rl_bh_get_search_slots.argtypes = [vp, C.c_uint64, _u32p, _u32p, _u32p]
    L.rl_bh_commit.argtypes = [vp, C.c_uint32, _u64p, _u64p]
    L.rl_bh_get_commit_result.argtypes = [vp, C.c_uint64, _u8p]
    L.rl_bh_update_connectivity.argtypes = [vp, C.c_uint32, C.POINTER(StepStats)]
    L.rl_bh_get_vacancies.argtypes = [vp, _u32p, _u32p, _u32p]
    L.rl_bh_tree_info.argtypes = [vp, _u64p, C.POINTER(C.c_uint32)]
    L.rl_bh_export_tree.argtypes = [vp, _u8p, _u8p, _u32p, _u8p, _f64p, _f64p, _f64p, _u32p, _u32p, _u8p, _u8p,
                                    _f64p, _f64p]
    L.rl_bh_last_search_ms.argtypes = [vp, _f64p]
    L.rl_bh_last_search_stats.argtypes = [vp, C.POINTER(StepStats)]
    L.rl_neurons_init.argtypes = [vp, C.POINTER(StepParams), _f64p, _f64p, _f64p, _f64p, _f64p, _u32p]
    L.rl_neurons_step.argtypes = [vp, C.c_uint32, _f64p]
    L.rl_neurons_commit_elements.argtypes = [vp, _u32p]
    L.rl_neurons_commit_elements_at.argtypes = [vp, C.c_uint32, _u32p]
    L.rl_neurons_get.argtypes = [vp, _f64p, _f64p, _f64p, _u8p, _f64p, _f64p, _u32p]
    L.rl_bh_graph_create.argtypes = [vp, C.c_uint32, C.c_uint64]
    L.rl_bh_graph_launch.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(StepStats)]
    L.rl_bh_graph_destroy.argtypes = [vp]
    L.rl_bh_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.rl_bh_candidate_lists.argtypes = [vp, C.c_uint64, _u32p, C.c_uint64, _u32p, _f64p, _u32p]
    L.rl_net_init.argtypes = [vp, C.c_uint64, _u64p, _u32p, C.POINTER(C.c_int32)]
    L.rl_net_get_edges.argtypes = [vp, C.c_uint64, _u64p, _u32p, C.POINTER(C.c_int32), _u64p]
    L.rl_bh_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.rl_bh_autotune.argtypes = [vp, C.c_uint32, C.POINTER(Tuning)]
    L.rl_bh_commit_requests_noalloc.argtypes = [C.c_uint64, C.c_uint64] + [vp] * 4 + [C.c_uint64, C.c_uint32] + \
        [vp] * 8 + [_u64p]
    for name in EXPORTED:
        f = getattr(L, name)
        if name not in ("rl_bh_last_error", "rl_bh_network_seed"):
            f.restype = C.c_int
    L.rl_bh_network_seed.restype = C.c_uint64
    _lib = L
    return L


class BHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class BarnesHutHIP:
    """Host mirror of RELeARN's BarnesHut algorithm object over the C ABI.

    Method names follow source/algorithm/Algorithm.h (init, prepare_update_connectivity,
    update_connectivity) and BarnesHut::find_target_neurons.
    """

    def __init__(self, box_min, box_max, theta=0.3, sigma=750.0, mu=0.0, seed=0,
                 precision=RL_PREC_FP64, device=0, flags=0):
        L = load()
        cfg = Config()
        for d in range(3):
            cfg.box_min[d] = float(box_min[d])
            cfg.box_max[d] = float(box_max[d])
        cfg.theta = theta
        cfg.sigma = sigma
        cfg.mu = mu
        cfg.seed = seed
        cfg.precision = precision
        cfg.device = device
        cfg.flags = flags
        h = C.c_void_p()
        st = L.rl_bh_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise BHError(st, L.rl_bh_last_error(None).decode())
        self.h = h
        self.L = L
        self.n = 0

    def _check(self, st):
        if st != 0:
            raise BHError(st, self.L.rl_bh_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.rl_bh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # Algorithm::init
    def init(self, xyz, signal):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        self.n = xyz.shape[0]
        self._check(self.L.rl_bh_init(self.h, self.n, _ptr(xyz, _f64p), _ptr(sig, _u8p)))

    def init_networks(self, offsets, boxes, xyz, signal):
        """Ensemble of independent networks: network g owns neurons offsets[g]:offsets[g+1] and the
        box boxes[g] = (min x, min y, min z, max x, max y, max z)."""
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        bx = np.ascontiguousarray(boxes, dtype=np.float64).reshape(-1, 6)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        self.n = xyz.shape[0]
        self._check(self.L.rl_bh_init_networks(self.h, len(off) - 1, _ptr(off, _u64p), _ptr(bx, _f64p),
                                               _ptr(xyz, _f64p), _ptr(sig, _u8p)))

    # Algorithm::create_neurons
    def create_neurons(self, xyz, signal, x=None, u=None, ca=None, target_ca=None, grown=None):
        """Appends neurons (ids n .. n + count - 1) at positions xyz with the given signal types; with
        neuron state, their initial x, u, calcium, target calcium and grown elements ((3, count))."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                for a in (x, u, ca, target_ca, grown)]
        self._check(self.L.rl_bh_create_neurons(self.h, xyz.shape[0], _ptr(xyz, _f64p), _ptr(sig, _u8p),
                                                *[_ptr(a, _f64p) for a in arrs]))
        self.n += xyz.shape[0]

    def set_enabled(self, enabled):
        en = None if enabled is None else np.ascontiguousarray(enabled, dtype=np.uint8)
        self._check(self.L.rl_bh_set_enabled(self.h, _ptr(en, _u8p)))

    # Algorithm::prepare_update_connectivity
    def prepare_update_connectivity(self, vac_axon, vac_exc, vac_inh):
        va = np.ascontiguousarray(vac_axon, dtype=np.uint32)
        ve = np.ascontiguousarray(vac_exc, dtype=np.uint32)
        vi = np.ascontiguousarray(vac_inh, dtype=np.uint32)
        self._check(self.L.rl_bh_prepare_update_connectivity(self.h, _ptr(va, _u32p), _ptr(ve, _u32p),
                                                             _ptr(vi, _u32p)))

    def prepare_update_connectivity_device(self, d_va, d_ve, d_vi):
        """Vacancy spans given as device addresses (ints) on the context's GPU."""
        self._check(self.L.rl_bh_prepare_update_connectivity_device(self.h, C.c_void_p(d_va), C.c_void_p(d_ve),
                                                                    C.c_void_p(d_vi)))

    def prepare_resident(self):
        self._check(self.L.rl_bh_prepare_resident(self.h))

    # BarnesHut::find_target_neurons
    def find_target_neurons(self, step=0):
        nreq = C.c_uint64()
        self._check(self.L.rl_bh_find_target_neurons(self.h, step, C.byref(nreq)))
        return int(nreq.value)

    def get_requests(self, m):
        t = np.zeros(max(m, 1), np.uint32)
        s = np.zeros(max(m, 1), np.uint32)
        g = np.zeros(max(m, 1), np.uint8)
        j = np.zeros(max(m, 1), np.uint32)
        self._check(self.L.rl_bh_get_requests(self.h, max(m, 1), _ptr(t, _u32p), _ptr(s, _u32p), _ptr(g, _u8p),
                                              _ptr(j, _u32p)))
        return t[:m], s[:m], g[:m], j[:m]

    def get_search_slots(self, slots):
        t = np.zeros(max(slots, 1), np.uint32)
        tests = np.zeros(max(slots, 1), np.uint32)
        cnt = np.zeros(max(self.n, 1), np.uint32)
        self._check(self.L.rl_bh_get_search_slots(self.h, max(slots, 1), _ptr(t, _u32p), _ptr(tests, _u32p),
                                                  _ptr(cnt, _u32p)))
        return t[:slots], tests[:slots], cnt[:self.n]

    def commit(self, step=0, order=None):
        created = C.c_uint64()
        o = None if order is None else np.ascontiguousarray(order, dtype=np.uint64)
        self._check(self.L.rl_bh_commit(self.h, step, _ptr(o, _u64p), C.byref(created)))
        return int(created.value)

    def get_commit_result(self, m):
        ok = np.zeros(max(m, 1), np.uint8)
        self._check(self.L.rl_bh_get_commit_result(self.h, max(m, 1), _ptr(ok, _u8p)))
        return ok[:m]

    # Algorithm::update_connectivity
    def update_connectivity(self, step=0):
        st = StepStats()
        self._check(self.L.rl_bh_update_connectivity(self.h, step, C.byref(st)))
        return st.as_dict()

    def net_init(self, capacity, in_offsets=None, sources=None, weights=None):
        """Network graph of the synaptic input: CSR by target (n + 1 offsets), later commits append."""
        off = None if in_offsets is None else np.ascontiguousarray(in_offsets, dtype=np.uint64)
        src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int32)
        self._check(self.L.rl_net_init(self.h, int(capacity), _ptr(off, _u64p), _ptr(src, _u32p),
                                       _ptr(w, C.POINTER(C.c_int32))))

    def net_get_edges(self):
        ne = C.c_uint64()
        self._check(self.L.rl_net_get_edges(self.h, 0, None, None, None, C.byref(ne)))
        m = int(ne.value)
        off = np.zeros(self.n + 1, np.uint64)
        src = np.zeros(max(m, 1), np.uint32)
        w = np.zeros(max(m, 1), np.int32)
        self._check(self.L.rl_net_get_edges(self.h, m, _ptr(off, _u64p), _ptr(src, _u32p),
                                            _ptr(w, C.POINTER(C.c_int32)), C.byref(ne)))
        return off, src[:m], w[:m]

    def candidate_lists(self, sources, capacity=4096):
        """Level-1 candidate lists (nodes, probabilities) of the given sources after a prepare."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        k = len(src)
        nodes = np.zeros(max(k * capacity, 1), np.uint32)
        probs = np.zeros(max(k * capacity, 1))
        counts = np.zeros(max(k, 1), np.uint32)
        self._check(self.L.rl_bh_candidate_lists(self.h, k, _ptr(src, _u32p), capacity, _ptr(nodes, _u32p),
                                                 _ptr(probs, _f64p), _ptr(counts, _u32p)))
        out = []
        for i in range(k):
            c = min(int(counts[i]), capacity)
            out.append((nodes[i * capacity:i * capacity + c].copy(), probs[i * capacity:i * capacity + c].copy(),
                        int(counts[i])))
        return out

    def set_tuning(self, search_waves_per_cu, search_chunk=1, walk_min_sources=0):
        """walk_min_sources: FP32-mixed, active sources from which an update takes the group walk
        (0 keeps the current value)."""
        t = Tuning()
        t.search_waves_per_cu = int(search_waves_per_cu)
        t.search_chunk = int(search_chunk)
        t.walk_min_sources = int(walk_min_sources)
        self._check(self.L.rl_bh_set_tuning(self.h, C.byref(t)))

    @staticmethod
    def _tuning_dict(t):
        return {"search_waves_per_cu": t.search_waves_per_cu, "search_chunk": t.search_chunk,
                "walk_min_sources": t.walk_min_sources}

    def get_tuning(self):
        t = Tuning()
        self._check(self.L.rl_bh_get_tuning(self.h, C.byref(t)))
        return self._tuning_dict(t)

    def autotune(self, step=0):
        """Keeps the fastest search tuning for the prepared update; returns it."""
        t = Tuning()
        self._check(self.L.rl_bh_autotune(self.h, int(step), C.byref(t)))
        return self._tuning_dict(t)

    def graph_create(self, steps=0, max_vacant_axons=None):
        """Captures a plasticity period (steps neuron steps, element commit, connectivity update)."""
        if max_vacant_axons is None:
            max_vacant_axons = 1
        self._check(self.L.rl_bh_graph_create(self.h, int(steps), int(max_vacant_axons)))

    def graph_launch(self, first_step=0, update=0, stats=True):
        """Replays the captured period; returns its stats dict, or None for an asynchronous replay."""
        if not stats:
            self._check(self.L.rl_bh_graph_launch(self.h, int(first_step), int(update), None))
            return None
        st = StepStats()
        self._check(self.L.rl_bh_graph_launch(self.h, int(first_step), int(update), C.byref(st)))
        return st.as_dict()

    def graph_destroy(self):
        self._check(self.L.rl_bh_graph_destroy(self.h))

    def get_vacancies(self):
        va = np.zeros(max(self.n, 1), np.uint32)
        ve = np.zeros(max(self.n, 1), np.uint32)
        vi = np.zeros(max(self.n, 1), np.uint32)
        self._check(self.L.rl_bh_get_vacancies(self.h, _ptr(va, _u32p), _ptr(ve, _u32p), _ptr(vi, _u32p)))
        return va[:self.n], ve[:self.n], vi[:self.n]

    def tree_info(self):
        m = C.c_uint64()
        d = C.c_uint32()
        self._check(self.L.rl_bh_tree_info(self.h, C.byref(m), C.byref(d)))
        return int(m.value), int(d.value)

    def export_tree(self):
        m, _ = self.tree_info()
        out = {
            "level": np.zeros(m, np.uint8), "nchild": np.zeros(m, np.uint8), "link": np.zeros(m, np.uint32),
            "octant": np.zeros(m, np.uint8), "box_lo": np.zeros((m, 3)), "box_hi": np.zeros((m, 3)),
            "side": np.zeros(m), "cnt_exc": np.zeros(m, np.uint32), "cnt_inh": np.zeros(m, np.uint32),
            "has_exc": np.zeros(m, np.uint8), "has_inh": np.zeros(m, np.uint8), "pos_exc": np.zeros((m, 3)),
            "pos_inh": np.zeros((m, 3)),
        }
        self._check(self.L.rl_bh_export_tree(
            self.h, _ptr(out["level"], _u8p), _ptr(out["nchild"], _u8p), _ptr(out["link"], _u32p),
            _ptr(out["octant"], _u8p), _ptr(out["box_lo"], _f64p), _ptr(out["box_hi"], _f64p),
            _ptr(out["side"], _f64p), _ptr(out["cnt_exc"], _u32p), _ptr(out["cnt_inh"], _u32p),
            _ptr(out["has_exc"], _u8p), _ptr(out["has_inh"], _u8p), _ptr(out["pos_exc"], _f64p),
            _ptr(out["pos_inh"], _f64p)))
        return out

    def last_search_ms(self):
        ms = C.c_double()
        self._check(self.L.rl_bh_last_search_ms(self.h, C.byref(ms)))
        return ms.value

    def last_search_stats(self):
        st = StepStats()
        self._check(self.L.rl_bh_last_search_stats(self.h, C.byref(st)))
        return st.as_dict()

    # neuron step
    def neurons_init(self, params: StepParams, target_ca, x=None, u=None, ca=None, grown=None, connected=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, u, ca, target_ca, grown)]
        con = None if connected is None else np.ascontiguousarray(connected, dtype=np.uint32)
        self._check(self.L.rl_neurons_init(self.h, C.byref(params), *[_ptr(a, _f64p) for a in arrs], _ptr(con, _u32p)))

    def neurons_set_network_growth(self, nu):
        a = None if nu is None else np.ascontiguousarray(nu, dtype=np.float64).ravel()
        self._check(self.L.rl_neurons_set_network_growth(self.h, _ptr(a, _f64p)))

    def neurons_set_element_params(self, min_calcium=None, retract_ratio=None):
        """Per-neuron minimum calcium and vacant retract ratio, (3, n) type-major each, or None for the
        per-type values of the step parameters."""
        mc = None if min_calcium is None else np.ascontiguousarray(min_calcium, dtype=np.float64).ravel()
        rr = None if retract_ratio is None else np.ascontiguousarray(retract_ratio, dtype=np.float64).ravel()
        self._check(self.L.rl_neurons_set_element_params(self.h, _ptr(mc, _f64p), _ptr(rr, _f64p)))

    def neurons_step(self, step, inp=None):
        i = None if inp is None else np.ascontiguousarray(inp, dtype=np.float64)
        self._check(self.L.rl_neurons_step(self.h, step, _ptr(i, _f64p)))

    def neurons_commit_elements(self, update=0):
        """Element commit of the three types (with a network graph, also the synapse deletions drawn
        for plasticity update `update`); returns the unbound elements per type and neuron."""
        d = np.zeros(max(3 * self.n, 1), np.uint32)
        self._check(self.L.rl_neurons_commit_elements_at(self.h, int(update), _ptr(d, _u32p)))
        return d[:3 * self.n].reshape(3, self.n)

    def neurons_get(self):
        n = self.n
        out = {"x": np.zeros(n), "u": np.zeros(n), "ca": np.zeros(n), "fired": np.zeros(n, np.uint8),
               "delta": np.zeros(3 * n), "grown": np.zeros(3 * n), "connected": np.zeros(3 * n, np.uint32)}
        self._check(self.L.rl_neurons_get(self.h, _ptr(out["x"], _f64p), _ptr(out["u"], _f64p),
                                          _ptr(out["ca"], _f64p), _ptr(out["fired"], _u8p),
                                          _ptr(out["delta"], _f64p), _ptr(out["grown"], _f64p),
                                          _ptr(out["connected"], _u32p)))
        out["delta"] = out["delta"].reshape(3, n)
        out["grown"] = out["grown"].reshape(3, n)
        out["connected"] = out["connected"].reshape(3, n)
        return out


def network_seed(seed, network):
    """Draw-stream seed of network `network` of an ensemble."""
    return int(load().rl_bh_network_seed(seed, network))


def commit_requests_noalloc(n, m, target, source, signal, axon_ordinal, network_seed, step, vacant, connected, ok,
                            stream=0, wait=True):
    """rl_bh_commit_requests_noalloc on device buffers given as integer addresses: target, source, signal,
    axon_ordinal (m requests), vacant = (axon, excitatory, inhibitory) vacancy arrays and connected = the
    same three connected-count arrays or None (n entries each), ok (m flags).  Returns the number of
    accepted requests when wait is true, else None (the call then does not wait for the device)."""
    L = load()
    created = C.c_uint64()
    conn = list(connected) if connected is not None else [None, None, None]
    st = L.rl_bh_commit_requests_noalloc(n, m, target, source, signal, axon_ordinal, network_seed, step,
                                         vacant[0], vacant[1], vacant[2], conn[0], conn[1], conn[2], ok,
                                         stream or None, C.byref(created) if wait else None)
    if st != 0:
        raise BHError(st, L.rl_bh_last_error(None).decode())
    return int(created.value) if wait else None
