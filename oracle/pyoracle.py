"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.  It loads
``oracle/libbh_oracle.so`` built by ``oracle/build.py``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbh_oracle.so")

STREAM_COUNTER = 0
STREAM_MT19937 = 1

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)


class SearchParams(C.Structure):
    _fields_ = [
        ("theta", C.c_double),
        ("sigma", C.c_double),
        ("mu", C.c_double),
        ("seed", C.c_uint64),
        ("step", C.c_uint32),
        ("stream", C.c_int32),
        ("threads", C.c_int32),
        ("reserved", C.c_int32),
    ]


class SearchStats(C.Structure):
    _fields_ = [
        ("searches", C.c_uint64),
        ("found", C.c_uint64),
        ("node_tests", C.c_uint64),
        ("node_tests_l1", C.c_uint64),
        ("candidates", C.c_uint64),
        ("picks", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("near_boundary", C.c_uint64),
        ("seconds", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


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
        ("reserved", C.c_uint32),
        ("tau_ca", C.c_double),
        ("beta_ca", C.c_double),
        ("eta", C.c_double * 3),
        ("nu", C.c_double * 3),
    ]


def _ptr(a, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle/build.py)")
        L = C.CDLL(LIB_PATH)
        L.orc_tree_build.argtypes = [C.c_uint64, _f64p, _f64p, _f64p, C.POINTER(C.c_void_p)]
        L.orc_tree_build.restype = C.c_int
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_free.restype = None
        L.orc_tree_num_nodes.argtypes = [C.c_void_p]
        L.orc_tree_num_nodes.restype = C.c_uint64
        L.orc_tree_depth.argtypes = [C.c_void_p]
        L.orc_tree_depth.restype = C.c_uint32
        L.orc_tree_update.argtypes = [C.c_void_p, _u32p, _u32p, _u8p, C.c_int]
        L.orc_tree_update.restype = C.c_int
        L.orc_tree_export.argtypes = [C.c_void_p, _u8p, _u8p, _u32p, _u8p, _f64p, _f64p, _f64p,
                                      _u32p, _u32p, _u8p, _u8p, _f64p, _f64p]
        L.orc_tree_export.restype = C.c_int
        L.orc_find_targets.argtypes = [C.c_void_p, C.c_uint64, _f64p, _u8p, _u32p, _u8p, C.c_uint64, C.c_uint64,
                                       _u64p, C.POINTER(SearchParams), _u32p, _u32p, _f64p, _u32p,
                                       C.POINTER(SearchStats)]
        L.orc_find_targets.restype = C.c_int
        L.orc_find_targets_faithful.argtypes = [C.c_void_p, C.c_uint64, _f64p, _u8p, _u32p, _u8p, _u64p, C.c_uint64,
                                                C.POINTER(SearchParams), _u64p, C.POINTER(SearchStats)]
        L.orc_find_targets_faithful.restype = C.c_int
        L.orc_candidate_list.argtypes = [C.c_void_p, C.c_uint64, _f64p, C.c_uint32, C.c_uint8, C.c_double, C.c_double,
                                         C.c_double, C.c_uint64, _u64p, _f64p, _u64p, _f64p, C.POINTER(C.c_int),
                                         _u64p]
        L.orc_candidate_list.restype = C.c_int
        L.orc_pick_index.argtypes = [_f64p, C.c_uint64, C.c_double]
        L.orc_pick_index.restype = C.c_int64
        L.orc_draw_to_uniform.argtypes = [C.c_uint32, C.c_double]
        L.orc_draw_to_uniform.restype = C.c_double
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_philox4x32_10.restype = None
        L.orc_pick_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_pick_draw.restype = C.c_uint32
        L.orc_commit_priority.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_commit_priority.restype = C.c_uint64
        L.orc_mt19937_kth.argtypes = [C.c_uint32, C.c_uint64]
        L.orc_mt19937_kth.restype = C.c_uint32
        L.orc_hypot3.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_hypot3.restype = C.c_double
        L.orc_commit.argtypes = [C.c_uint64, _u32p, _u32p, _u8p, _u32p, _u64p, C.c_uint64, C.c_uint32,
                                 _u32p, _u32p, _u32p, _u32p, _u32p, _u32p, _u8p, _u64p]
        L.orc_commit.restype = C.c_int
        L.orc_reference_shuffle.argtypes = [C.c_uint64, C.c_uint32, _u64p]
        L.orc_reference_shuffle.restype = C.c_int
        L.orc_neuron_step.argtypes = [C.c_uint64, C.POINTER(StepParams), _f64p, _f64p, _f64p, _f64p, _u8p, _f64p,
                                      _f64p, _u8p, _f64p]
        L.orc_neuron_step.restype = C.c_int
        L.orc_graph_create.argtypes = [C.c_uint64, _u64p, _u32p, C.POINTER(C.c_int32), C.POINTER(C.c_void_p)]
        L.orc_graph_create.restype = C.c_int
        L.orc_graph_free.argtypes = [C.c_void_p]
        L.orc_graph_free.restype = None
        L.orc_graph_add.argtypes = [C.c_void_p, C.c_uint64, _u32p, _u32p, C.POINTER(C.c_int32)]
        L.orc_graph_add.restype = C.c_int
        L.orc_graph_num_edges.argtypes = [C.c_void_p]
        L.orc_graph_num_edges.restype = C.c_uint64
        L.orc_graph_export.argtypes = [C.c_void_p, C.c_uint64, _u64p, _u32p, C.POINTER(C.c_int32)]
        L.orc_graph_export.restype = C.c_int
        L.orc_delete_synapses.argtypes = [C.c_void_p, C.c_int, _u32p, _u32p, _u64p, _u64p, C.c_uint32,
                                          _u32p, _u32p, _u32p, _u32p, _u32p, _u32p, _u8p, _u64p]
        L.orc_delete_synapses.restype = C.c_int
        L.orc_deletion_key.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_deletion_key.restype = C.c_uint64
        L.orc_synaptic_input.argtypes = [C.c_uint64, _u64p, _u32p, C.POINTER(C.c_int32), _u8p, C.c_double, _f64p,
                                         _f64p]
        L.orc_synaptic_input.restype = C.c_int
        L.orc_elements_update.argtypes = [C.c_uint64, _f64p, _f64p, _u32p, _u32p, _f64p, _u32p]
        L.orc_elements_update.restype = C.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what} failed with status {st}")


class Tree:
    """Pointer octree of the oracle (OctreeNode::insert by serial insertion)."""

    def __init__(self, xyz, box_min, box_max):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.n = xyz.shape[0]
        bmin = np.ascontiguousarray(box_min, dtype=np.float64)
        bmax = np.ascontiguousarray(box_max, dtype=np.float64)
        h = C.c_void_p()
        st = lib().orc_tree_build(self.n, _ptr(xyz, _f64p), _ptr(bmin, _f64p), _ptr(bmax, _f64p), C.byref(h))
        self.status = st
        _check(st, "orc_tree_build")
        self.h = h
        self.xyz = xyz

    def close(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_nodes(self):
        return int(lib().orc_tree_num_nodes(self.h))

    @property
    def depth(self):
        return int(lib().orc_tree_depth(self.h))

    def update(self, vac_exc, vac_inh, enabled=None, threads=0):
        ve = np.ascontiguousarray(vac_exc, dtype=np.uint32)
        vi = np.ascontiguousarray(vac_inh, dtype=np.uint32)
        en = None if enabled is None else np.ascontiguousarray(enabled, dtype=np.uint8)
        _check(lib().orc_tree_update(self.h, _ptr(ve, _u32p), _ptr(vi, _u32p), _ptr(en, _u8p), threads),
               "orc_tree_update")

    def export(self):
        m = self.num_nodes
        out = {
            "level": np.zeros(m, np.uint8),
            "nchild": np.zeros(m, np.uint8),
            "link": np.zeros(m, np.uint32),
            "octant": np.zeros(m, np.uint8),
            "box_lo": np.zeros((m, 3), np.float64),
            "box_hi": np.zeros((m, 3), np.float64),
            "side": np.zeros(m, np.float64),
            "cnt_exc": np.zeros(m, np.uint32),
            "cnt_inh": np.zeros(m, np.uint32),
            "has_exc": np.zeros(m, np.uint8),
            "has_inh": np.zeros(m, np.uint8),
            "pos_exc": np.zeros((m, 3), np.float64),
            "pos_inh": np.zeros((m, 3), np.float64),
        }
        _check(lib().orc_tree_export(
            self.h, _ptr(out["level"], _u8p), _ptr(out["nchild"], _u8p), _ptr(out["link"], _u32p),
            _ptr(out["octant"], _u8p), _ptr(out["box_lo"], _f64p), _ptr(out["box_hi"], _f64p),
            _ptr(out["side"], _f64p), _ptr(out["cnt_exc"], _u32p), _ptr(out["cnt_inh"], _u32p),
            _ptr(out["has_exc"], _u8p), _ptr(out["has_inh"], _u8p), _ptr(out["pos_exc"], _f64p),
            _ptr(out["pos_inh"], _f64p)), "orc_tree_export")
        return out

    def candidate_list(self, root, src, source_id, sig, theta, sigma, mu=0.0, cap=1 << 20):
        srcv = np.ascontiguousarray(src, dtype=np.float64)
        nodes = np.zeros(cap, np.uint64)
        prob = np.zeros(cap, np.float64)
        ln = C.c_uint64()
        total = C.c_double()
        fb = C.c_int()
        tests = C.c_uint64()
        _check(lib().orc_candidate_list(self.h, root, _ptr(srcv, _f64p), source_id, sig, theta, sigma, mu, cap,
                                        _ptr(nodes, _u64p), _ptr(prob, _f64p), C.byref(ln), C.byref(total),
                                        C.byref(fb), C.byref(tests)), "orc_candidate_list")
        k = int(ln.value)
        return nodes[:k].copy(), prob[:k].copy(), total.value, bool(fb.value), int(tests.value)

    def find_targets(self, signal, vac_axon, enabled=None, *, theta=0.3, sigma=750.0, mu=0.0, seed=0, step=0,
                     stream=STREAM_COUNTER, threads=0, first=0, count=None):
        n = self.n
        if count is None:
            count = n - first
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        va = np.ascontiguousarray(vac_axon, dtype=np.uint32)
        en = None if enabled is None else np.ascontiguousarray(enabled, dtype=np.uint8)
        sub = va[first:first + count].astype(np.uint64)
        offs = np.zeros(count, np.uint64)
        if count > 1:
            offs[1:] = np.cumsum(sub)[:-1]
        total_slots = int(sub.sum())
        tgt = np.full(max(total_slots, 1), 0xFFFFFFFF, np.uint32)
        cnt = np.zeros(count, np.uint32)
        margin = np.full(max(total_slots, 1), np.inf, np.float64)
        tests = np.zeros(max(total_slots, 1), np.uint32)
        p = SearchParams(theta, sigma, mu, seed, step, stream, threads, 0)
        st = SearchStats()
        _check(lib().orc_find_targets(self.h, n, _ptr(self.xyz, _f64p), _ptr(sig, _u8p), _ptr(va, _u32p),
                                      _ptr(en, _u8p), first, count, _ptr(offs, _u64p), C.byref(p),
                                      _ptr(tgt, _u32p), _ptr(cnt, _u32p), _ptr(margin, _f64p), _ptr(tests, _u32p),
                                      C.byref(st)), "orc_find_targets")
        return {"offsets": offs, "target": tgt[:total_slots], "count": cnt, "margin": margin[:total_slots],
                "node_tests": tests[:total_slots], "stats": st.as_dict()}

    def find_targets_faithful(self, signal, vac_axon, sources, enabled=None, *, theta=0.3, sigma=750.0, mu=0.0,
                              seed=0, step=0, stream=STREAM_MT19937, threads=0):
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        va = np.ascontiguousarray(vac_axon, dtype=np.uint32)
        en = None if enabled is None else np.ascontiguousarray(enabled, dtype=np.uint8)
        src = np.ascontiguousarray(sources, dtype=np.uint64)
        p = SearchParams(theta, sigma, mu, seed, step, stream, threads, 0)
        st = SearchStats()
        nreq = C.c_uint64()
        _check(lib().orc_find_targets_faithful(self.h, self.n, _ptr(self.xyz, _f64p), _ptr(sig, _u8p),
                                               _ptr(va, _u32p), _ptr(en, _u8p), _ptr(src, _u64p), len(src),
                                               C.byref(p), C.byref(nreq), C.byref(st)), "orc_find_targets_faithful")
        d = st.as_dict()
        d["requests"] = int(nreq.value)
        return d


def pick_index(prob, r):
    p = np.ascontiguousarray(prob, dtype=np.float64)
    return int(lib().orc_pick_index(_ptr(p, _f64p), len(p), float(r)))


def draw_to_uniform(k, next_total):
    return float(lib().orc_draw_to_uniform(int(k), float(next_total)))


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
    return out


def pick_draw(seed, neuron, step, j, level):
    return int(lib().orc_pick_draw(seed, neuron, step, j, level))


def commit_priority(seed, source, step, j):
    return int(lib().orc_commit_priority(seed, source, step, j))


def mt19937_kth(seed, k):
    return int(lib().orc_mt19937_kth(seed, k))


def hypot3(x, y, z):
    return float(lib().orc_hypot3(x, y, z))


def commit(target, source, signal, req_j, vac_exc, vac_inh, vac_axon, *, order=None, seed=0, step=0,
           conn_exc=None, conn_inh=None, conn_axon=None):
    m = len(target)
    t = np.ascontiguousarray(target, dtype=np.uint32)
    s = np.ascontiguousarray(source, dtype=np.uint32)
    g = np.ascontiguousarray(signal, dtype=np.uint8)
    j = np.ascontiguousarray(req_j, dtype=np.uint32)
    o = None if order is None else np.ascontiguousarray(order, dtype=np.uint64)
    ok = np.zeros(max(m, 1), np.uint8)
    created = C.c_uint64()
    for a in (vac_exc, vac_inh, vac_axon):
        assert a.dtype == np.uint32 and a.flags.c_contiguous
    _check(lib().orc_commit(m, _ptr(t, _u32p), _ptr(s, _u32p), _ptr(g, _u8p), _ptr(j, _u32p), _ptr(o, _u64p),
                            seed, step, _ptr(vac_exc, _u32p), _ptr(vac_inh, _u32p), _ptr(vac_axon, _u32p),
                            _ptr(conn_exc, _u32p), _ptr(conn_inh, _u32p), _ptr(conn_axon, _u32p),
                            _ptr(ok, _u8p), C.byref(created)), "orc_commit")
    return ok[:m], int(created.value)


def reference_shuffle(m, mt_seed):
    perm = np.zeros(max(m, 1), np.uint64)
    _check(lib().orc_reference_shuffle(m, mt_seed, _ptr(perm, _u64p)), "orc_reference_shuffle")
    return perm[:m]


def neuron_step(params: StepParams, x, u, ca, inp, fired, target_ca, delta, enabled=None, min_calcium=None):
    n = len(x)
    en = None if enabled is None else np.ascontiguousarray(enabled, dtype=np.uint8)
    mc = None if min_calcium is None else np.ascontiguousarray(min_calcium, dtype=np.float64).ravel()
    assert mc is None or len(mc) == 3 * n
    for a in (x, u, ca, inp, target_ca, delta):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    assert fired.dtype == np.uint8
    _check(lib().orc_neuron_step(n, C.byref(params), _ptr(x, _f64p), _ptr(u, _f64p), _ptr(ca, _f64p),
                                 _ptr(inp, _f64p), _ptr(fired, _u8p), _ptr(target_ca, _f64p), _ptr(delta, _f64p),
                                 _ptr(en, _u8p), _ptr(mc, _f64p)), "orc_neuron_step")


class Graph:
    """Network graph restatement: pairs merged, zero weights erased, in- and out-edges per neuron."""

    def __init__(self, n, in_off=None, src=None, w=None):
        self.n = n
        off = None if in_off is None else np.ascontiguousarray(in_off, dtype=np.uint64)
        s = None if src is None else np.ascontiguousarray(src, dtype=np.uint32)
        ww = None if w is None else np.ascontiguousarray(w, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().orc_graph_create(n, _ptr(off, _u64p), _ptr(s, _u32p), _ptr(ww, C.POINTER(C.c_int32)),
                                      C.byref(h)), "orc_graph_create")
        self.h = h

    def add(self, target, source, weight):
        t = np.ascontiguousarray(target, dtype=np.uint32)
        s = np.ascontiguousarray(source, dtype=np.uint32)
        w = np.ascontiguousarray(weight, dtype=np.int32)
        _check(lib().orc_graph_add(self.h, len(t), _ptr(t, _u32p), _ptr(s, _u32p), _ptr(w, C.POINTER(C.c_int32))),
               "orc_graph_add")

    def export(self):
        m = int(lib().orc_graph_num_edges(self.h))
        off = np.zeros(self.n + 1, np.uint64)
        src = np.zeros(max(m, 1), np.uint32)
        w = np.zeros(max(m, 1), np.int32)
        _check(lib().orc_graph_export(self.h, m, _ptr(off, _u64p), _ptr(src, _u32p), _ptr(w, C.POINTER(C.c_int32))),
               "orc_graph_export")
        return off, src[:m], w[:m]

    def delete(self, etype, to_delete, signal, conn, vac, *, step, seed=0, net_of=None, net_off=None, net_seed=None):
        """Deletes the synapses of element type etype (0 axon, 1 exc, 2 inh dendrite); conn and vac are
        (3, n) uint32 arrays updated in place.  Returns the number of deleted units."""
        n = self.n
        td = np.ascontiguousarray(to_delete, dtype=np.uint32)
        no = np.zeros(n, np.uint32) if net_of is None else np.ascontiguousarray(net_of, dtype=np.uint32)
        nf = np.zeros(1, np.uint64) if net_off is None else np.ascontiguousarray(net_off, dtype=np.uint64)
        ns = np.array([seed], np.uint64) if net_seed is None else np.ascontiguousarray(net_seed, dtype=np.uint64)
        sig = np.ascontiguousarray(signal, dtype=np.uint8)
        assert conn.dtype == np.uint32 and vac.dtype == np.uint32 and conn.flags.c_contiguous and vac.flags.c_contiguous
        tot = C.c_uint64()
        _check(lib().orc_delete_synapses(self.h, int(etype), _ptr(td, _u32p), _ptr(no, _u32p), _ptr(nf, _u64p),
                                         _ptr(ns, _u64p), int(step), _ptr(conn[0], _u32p), _ptr(conn[1], _u32p),
                                         _ptr(conn[2], _u32p), _ptr(vac[0], _u32p), _ptr(vac[1], _u32p),
                                         _ptr(vac[2], _u32p), _ptr(sig, _u8p), C.byref(tot)), "orc_delete_synapses")
        return int(tot.value)

    def close(self):
        if getattr(self, "h", None):
            lib().orc_graph_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def deletion_key(seed, neuron, step, partner, etype, copy):
    return int(lib().orc_deletion_key(seed, neuron, step, partner, etype, copy))


def synaptic_input(in_off, in_src, in_w, fired, conductance, background=None):
    """Background plus equally weighted synaptic input of every neuron (in-edges CSR by target)."""
    n = len(in_off) - 1
    off = np.ascontiguousarray(in_off, dtype=np.uint64)
    src = np.ascontiguousarray(in_src, dtype=np.uint32)
    w = np.ascontiguousarray(in_w, dtype=np.int32)
    f = np.ascontiguousarray(fired, dtype=np.uint8)
    bg = None if background is None else np.ascontiguousarray(background, dtype=np.float64)
    out = np.zeros(max(n, 1))
    _check(lib().orc_synaptic_input(n, _ptr(off, _u64p), _ptr(src, _u32p), _ptr(w, C.POINTER(C.c_int32)),
                                    _ptr(f, _u8p), float(conductance), _ptr(bg, _f64p), _ptr(out, _f64p)),
           "orc_synaptic_input")
    return out[:n]


def elements_update(grown, delta, connected, vacant, retract_ratio=None):
    n = len(grown)
    out = np.zeros(max(n, 1), np.uint32)
    rr = None if retract_ratio is None else np.ascontiguousarray(retract_ratio, dtype=np.float64)
    _check(lib().orc_elements_update(n, _ptr(grown, _f64p), _ptr(delta, _f64p), _ptr(connected, _u32p),
                                     _ptr(vacant, _u32p), _ptr(rr, _f64p), _ptr(out, _u32p)), "orc_elements_update")
    return out[:n]
