// Barnes-Hut target search on the device.
//
// One lane per source neuron, sources in Morton order so that the lanes of a wavefront walk nearly
// the same part of the tree.  For every source the traversal reproduces
// BarnesHutBase::get_nodes_to_consider (source/algorithm/BarnesHutInternal/BarnesHutBase.h:129-206):
// a LIFO stack seeded with the children of the subtree root, acceptance test
// test_acceptance_criterion (:72-114), accepted nodes in stack order.  Children of a node are stored
// contiguously, so a stack entry is (first child, children left) and popping the last child first
// visits children in descending octant order exactly as the reference's pointer stack does.  The
// stack lives in LDS, one column per lane.
//
// The candidate list is never stored: pass 1 sums the probabilities (Kernel::create_probability_interval,
// source/algorithm/Kernel/Kernel.h:90-134, including the n/r fallback when the Gaussian sum
// underflows to 0), pass 2 walks the same list again and resolves up to eight draws at once with
// the picker's rule (ProbabilityPicker::pick_target, source/util/ProbabilityPicker.h:46-112): first
// candidate whose running sum reaches r.  The level-1 list depends only on the source, so all vacant
// axons of a neuron share it; every axon then descends from its picked node (find_target_neuron,
// BarnesHutBase.h:223-255) with its own draws, and a neuron stops at its first failed search
// (find_target_neurons, :283-289).
//
// Precision modes:
//   FP64        every distance with the libstdc++ three-argument hypot, the Gaussian with FP64 exp,
//               sums in FP64 in list order.
//   FP32_MIXED  distances from 31-bit fixed-point centroid differences in FP32, Gaussian with FP32
//               exp2; FP64 sums.  Acceptance decisions within a relative guard band of the threshold
//               and exact coincidences are re-decided with the FP64 test, so the candidate lists are
//               those of the FP64 path; a list whose FP32 sum is 0 is redone in FP64.
#include <hipcub/hipcub.hpp>

#include "rl_search_eval.h"

namespace rl {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

struct SumVisitor {
    double sum = 0.0;
    uint32_t count = 0;
    __device__ __forceinline__ void operator()(uint32_t, double p) {
        sum += p;
        count++;
    }
};

struct PickVisitor {
    double sum = 0.0;
    double r[kDrawsPerPass];
    uint32_t pick[kDrawsPerPass];
    int nd = 0;
    uint32_t last_positive = kNone;
    __device__ __forceinline__ void operator()(uint32_t node, double p) {
        const double s = sum + p;
        if (p > 0.0) {
#pragma unroll
            for (int k = 0; k < kDrawsPerPass; k++) {
                if (k < nd && pick[k] == kNone && s >= r[k]) {
                    pick[k] = node;
                }
            }
            last_positive = node;
        }
        sum = s;
    }
};

// Traversal of the subtree below `root` (get_nodes_to_consider); the visitor sees the accepted nodes
// in reference order with their probability.  Returns the number of acceptance tests.
template <bool kMixed, typename V>
__device__ uint32_t traverse(const SearchArgs& a, const Lane& L, uint32_t root, int kind, uint32_t* stk,
                             uint32_t stride, V& v) {
    const uint2* __restrict__ meta = L.meta;
    const uint2 mr = meta[root];
    if (mr.x == 0) {
        return 0;
    }
    uint32_t tests = 0;
    const uint32_t rnc = mr.y >> kLinkBits;
    if (rnc == 0) {
        tests++;
        const Eval e = kMixed ? eval_mixed(a, L, root, mr.x, true) : eval_exact(a, L, root, mr.x, true);
        if (e.status != kDiscard) {
            const bool self = (mr.y & kLinkMask) == L.src;
            v(root, self ? 0.0 : probability<kMixed>(a, L, e, mr.x, kind));
        }
        return tests;
    }
    // stack of (first child << 4 | children left); the top entry lives in a register and only the
    // entries below it are in LDS, so a pop touches LDS only when the top range is used up
    uint32_t top = ((mr.y & kLinkMask) << 4U) | rnc;
    int sp = 0;
    while (true) {
        if (top == 0) {
            if (sp == 0) {
                break;
            }
            sp--;
            top = stk[sp * stride];
        }
        const uint32_t left = (top & 15U) - 1U;
        const uint32_t node = (top >> 4U) + left;
        top = left == 0 ? 0U : ((top & ~15U) | left);
        tests++;
        // issue the count/link and the position loads together
        const uint2 m = meta[node];
        double4 r64 = make_double4(0.0, 0.0, 0.0, 0.0);
        uint4 r32 = make_uint4(0U, 0U, 0U, 0U);
        if (kMixed) {
            r32 = L.rec32[node];
            __asm__ volatile("" ::"v"(r32.x));  // keep the load ahead of the count test
        } else {
            r64 = L.rec64[node];
            __asm__ volatile("" ::"v"(r64.x));
        }
        if (m.x == 0) {
            continue;
        }
        const uint32_t nc = m.y >> kLinkBits;
        const bool leaf = nc == 0;
        const Eval e = kMixed ? eval_mixed_rec(a, L, r32, node, leaf) : eval_exact_rec(a, L, r64, leaf);
        if (e.status == kDiscard) {
            continue;
        }
        if (e.status == kAccept) {
            const bool self = leaf && (m.y & kLinkMask) == L.src;
            v(node, self ? 0.0 : probability<kMixed>(a, L, e, m.x, kind));
            continue;
        }
        if (top != 0) {
            stk[sp * stride] = top;
            sp++;
        }
        top = ((m.y & kLinkMask) << 4U) | nc;
    }
    return tests;
}

struct ListSum {
    double sum;
    uint32_t count;
    uint32_t tests_ref;    // tests of one reference traversal
    uint32_t tests_exec;   // tests executed here
    bool exact;            // list probabilities from the FP64 formulas
    bool redo;             // the Gaussian sum underflowed: the source goes to the fallback launch
    int kind;
};

// Pass 1 on a subtree: probability total.  The reference's fallbacks (FP64 re-evaluation of an
// FP32 sum of 0, then n/r when the Gaussian sum is 0) are compiled only into the fallback
// instantiation, which keeps the register footprint of the main launch small.
template <bool kMixed, bool kFallback>
__device__ ListSum list_sum(const SearchArgs& a, const Lane& L, uint32_t root, uint32_t* stk, uint32_t stride) {
    ListSum ls;
    SumVisitor v;
    ls.tests_ref = traverse<kMixed>(a, L, root, kGauss, stk, stride, v);
    ls.tests_exec = ls.tests_ref;
    ls.exact = !kMixed;
    ls.redo = false;
    ls.kind = kGauss;
    ls.sum = v.sum;
    ls.count = v.count;
    if (v.count > 0 && v.sum == 0.0) {
        if (!kFallback) {
            ls.redo = true;
            return ls;
        }
        if (kMixed) {
            SumVisitor w;
            ls.tests_exec += traverse<false>(a, L, root, kGauss, stk, stride, w);
            ls.exact = true;
            ls.sum = w.sum;
        }
        if (ls.sum == 0.0) {
            SumVisitor w;
            ls.tests_exec += traverse<false>(a, L, root, kInverse, stk, stride, w);
            ls.exact = true;
            ls.kind = kInverse;
            ls.sum = w.sum;
        }
    }
    return ls;
}

template <bool kMixed, bool kFallback>
__device__ uint32_t list_pick(const SearchArgs& a, const Lane& L, uint32_t root, const ListSum& ls, PickVisitor& pv,
                              uint32_t* stk, uint32_t stride) {
    uint32_t t;
    if (kFallback && ls.exact) {
        t = traverse<false>(a, L, root, ls.kind, stk, stride, pv);
    } else {
        t = traverse<kMixed>(a, L, root, ls.kind, stk, stride, pv);
    }
#pragma unroll
    for (int k = 0; k < kDrawsPerPass; k++) {
        if (k < pv.nd && pv.pick[k] == kNone) {
            pv.pick[k] = pv.last_positive;  // r above the recomputed sum: last positive candidate
        }
    }
    return t;
}

template <bool kMixed, bool kFallback>
__global__ void __launch_bounds__(kBlock) k_search(SearchArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= a.nwork) {
        return;
    }
    uint32_t* stk = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;
    const uint32_t src = a.order[t];
    const uint32_t na = a.vac_axon[src];
    const uint64_t slot0 = a.slot_off[src];
    if (a.enabled[src] == 0 || na == 0) {
        a.found[src] = 0;
        a.searched[src] = 0;
        return;
    }
    Lane L;
    lane_source(a, L, src, kMixed);
    uint64_t exec = 0;
    uint64_t ref = 0;
    const uint2 mroot = L.meta[L.root];
    if ((mroot.y >> kLinkBits) == 0) {
        // the root is a leaf (one neuron): it holds the source, so no target (BarnesHutBase.h:231-237)
        const uint32_t nid = mroot.y & kLinkMask;
        uint32_t f = 0;
        if (nid != src) {
            for (uint32_t j = 0; j < na; j++) {
                a.slot_target[slot0 + j] = nid;
                a.slot_tests[slot0 + j] = 0;
            }
            f = na;
        } else {
            a.slot_tests[slot0] = 0;
        }
        a.found[src] = f;
        a.searched[src] = f == na ? na : 1;
        atomicAdd(&a.counters[2], static_cast<unsigned long long>(f == na ? na : 1));
        return;
    }
    // level 1: one list for all axons of the neuron
    const ListSum l1 = list_sum<kMixed, kFallback>(a, L, L.root, stk, stride);
    if (l1.redo) {
        a.redo[atomicAdd(a.n_redo, 1U)] = src;
        return;
    }
    exec += l1.tests_exec;
    if (l1.count == 0 || l1.sum == 0.0) {
        a.slot_tests[slot0] = l1.tests_ref;
        a.found[src] = 0;
        a.searched[src] = 1;
        atomicAdd(&a.counters[0], static_cast<unsigned long long>(exec));
        atomicAdd(&a.counters[1], static_cast<unsigned long long>(l1.tests_ref));
        atomicAdd(&a.counters[2], 1ULL);
        return;
    }
    const double next1 = nextafter(l1.sum, l1.sum + 0.00001);
    for (uint32_t j0 = 0; j0 < na; j0 += kDrawsPerPass) {
        PickVisitor pv;
        pv.nd = static_cast<int>(min(na - j0, static_cast<uint32_t>(kDrawsPerPass)));
#pragma unroll
        for (int k = 0; k < kDrawsPerPass; k++) {
            pv.pick[k] = kNone;
            pv.r[k] = k < pv.nd ? draw_to_uniform(pick_draw(L.seed, L.local, a.step, j0 + k, 1U), next1) : 0.0;
        }
        exec += list_pick<kMixed, kFallback>(a, L, L.root, l1, pv, stk, stride);
#pragma unroll
        for (int k = 0; k < kDrawsPerPass; k++) {
            if (k < pv.nd) {
                a.slot_target[slot0 + j0 + k] = pv.pick[k];
            }
        }
    }
    // descents, axon by axon, stopping at the first failure
    uint32_t found = 0;
    uint32_t searched = 0;
    for (uint32_t j = 0; j < na; j++) {
        uint32_t node = a.slot_target[slot0 + j];
        uint32_t tests_j = l1.tests_ref;
        bool ok = true;
        searched++;
        for (uint32_t level = 2;; level++) {
            if (node >= a.m || level > kMaxLevels + 2) {
                ok = false;  // unreachable for a consistent tree; keeps every access in bounds
                break;
            }
            const uint2 m = L.meta[node];
            if ((m.y >> kLinkBits) == 0) {
                node = m.y & kLinkMask;  // leaf: the target neuron
                break;
            }
            const ListSum ls = list_sum<kMixed, kFallback>(a, L, node, stk, stride);
            if (ls.redo) {
                a.redo[atomicAdd(a.n_redo, 1U)] = src;
                return;
            }
            exec += ls.tests_exec;
            tests_j += ls.tests_ref;
            if (ls.count == 0 || ls.sum == 0.0) {
                ok = false;
                break;
            }
            PickVisitor pv;
            pv.nd = 1;
#pragma unroll
            for (int k = 0; k < kDrawsPerPass; k++) {
                pv.pick[k] = kNone;
                pv.r[k] = 0.0;
            }
            pv.r[0] = draw_to_uniform(pick_draw(L.seed, L.local, a.step, j, level), nextafter(ls.sum, ls.sum + 0.00001));
            exec += list_pick<kMixed, kFallback>(a, L, node, ls, pv, stk, stride);
            node = pv.pick[0];
        }
        a.slot_tests[slot0 + j] = tests_j;
        ref += tests_j;
        if (!ok) {
            a.slot_target[slot0 + j] = kNone;
            break;
        }
        a.slot_target[slot0 + j] = node;
        found++;
    }
    a.found[src] = found;
    a.searched[src] = searched;
    atomicAdd(&a.counters[0], static_cast<unsigned long long>(exec));
    atomicAdd(&a.counters[1], static_cast<unsigned long long>(ref));
    atomicAdd(&a.counters[2], static_cast<unsigned long long>(searched));
}

// Descents that the cooperative kernels hand over (a level that needs expansion or the reference's
// fallbacks): one lane continues the search of one slot from (node, level) with the general
// traversal, as the descent loop of k_search does.
template <bool kMixed>
__global__ void __launch_bounds__(kBlock) k_descend_lane(SearchArgs a, const uint4* __restrict__ items, uint64_t count) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= count) {
        return;
    }
    uint32_t* stk = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;
    const uint4 it = items[t];
    const uint32_t slot = it.x, src = it.y;
    uint32_t node = it.z;
    Lane L;
    lane_source(a, L, src, kMixed);
    const uint32_t j = slot - static_cast<uint32_t>(a.slot_off[src]);
    uint64_t exec = 0;
    uint32_t tests = 0;
    bool ok = true;
    for (uint32_t level = it.w;; level++) {
        if (node >= a.m || level > kMaxLevels + 2) {
            ok = false;
            break;
        }
        const uint2 m = L.meta[node];
        if ((m.y >> kLinkBits) == 0) {
            node = m.y & kLinkMask;
            break;
        }
        const ListSum ls = list_sum<kMixed, true>(a, L, node, stk, stride);
        exec += ls.tests_exec;
        tests += ls.tests_ref;
        if (ls.count == 0 || ls.sum == 0.0) {
            ok = false;
            break;
        }
        PickVisitor pv;
        pv.nd = 1;
#pragma unroll
        for (int k = 0; k < kDrawsPerPass; k++) {
            pv.pick[k] = kNone;
            pv.r[k] = 0.0;
        }
        pv.r[0] = draw_to_uniform(pick_draw(L.seed, L.local, a.step, j, level), nextafter(ls.sum, ls.sum + 0.00001));
        exec += list_pick<kMixed, true>(a, L, node, ls, pv, stk, stride);
        node = pv.pick[0];
    }
    a.slot_tests[slot] += tests;
    a.slot_target[slot] = ok ? node : kNone;
    atomicAdd(&a.counters[0], static_cast<unsigned long long>(exec));
}

__global__ void k_widen(uint64_t n, const uint32_t* __restrict__ v, uint64_t* __restrict__ out) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        out[i] = v[i];
    } else if (i == n) {
        out[i] = 0;
    }
}

__global__ void k_emit_requests(uint64_t n, const uint32_t* __restrict__ found, const uint64_t* __restrict__ req_off,
                                const uint64_t* __restrict__ slot_off, const uint32_t* __restrict__ slot_target,
                                const uint8_t* __restrict__ signal, uint32_t* __restrict__ rt, uint32_t* __restrict__ rs,
                                uint8_t* __restrict__ rg, uint32_t* __restrict__ rj) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t f = found[i];
    const uint64_t o = req_off[i];
    const uint64_t s = slot_off[i];
    for (uint32_t j = 0; j < f; j++) {
        rt[o + j] = slot_target[s + j];
        rs[o + j] = static_cast<uint32_t>(i);
        rg[o + j] = signal[i];
        rj[o + j] = j;
    }
}

__global__ void k_active_flags(uint64_t n, const uint32_t* __restrict__ order, const uint8_t* __restrict__ enabled,
                               const uint32_t* __restrict__ vac, uint32_t* __restrict__ flag) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t < n) {
        const uint32_t src = order[t];
        flag[t] = (enabled[src] != 0 && vac[src] != 0) ? 1U : 0U;
    }
}

}  // namespace

int launch_descend_lane(rl_bh_ctx* c, const SearchArgs& sa, const uint4* items, uint64_t count) {
    hipStream_t s = c->stream;
    const size_t lds = static_cast<size_t>(std::max<uint32_t>(c->tree.depth, 1U)) * kBlock * sizeof(uint32_t);
    if (c->cfg.precision == RL_PREC_FP32_MIXED) {
        k_descend_lane<true><<<grid_for(count), kBlock, lds, s>>>(sa, items, count);
    } else {
        k_descend_lane<false><<<grid_for(count), kBlock, lds, s>>>(sa, items, count);
    }
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

// Active sources (enabled, vacant axons > 0) in Morton order into c->coop_work.
int compact_active(rl_bh_ctx* c, uint64_t* nwork) {
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    k_active_flags<<<grid_for(n), kBlock, 0, s>>>(n, c->order, c->enabled, c->vac[0], c->coop_flag);
    RL_HIP(c, hipGetLastError());
    size_t tb = 0;
    RL_HIP(c, hipcub::DeviceSelect::Flagged(nullptr, tb, c->order, c->coop_flag, c->coop_work, c->coop_nwork,
                                            static_cast<int>(n), s));
    RL_TRY(ensure_tmp(c, tb));
    RL_HIP(c, hipcub::DeviceSelect::Flagged(c->tmp, tb, c->order, c->coop_flag, c->coop_work, c->coop_nwork,
                                            static_cast<int>(n), s));
    int nw = 0;
    RL_HIP(c, hipMemcpyAsync(&nw, c->coop_nwork, sizeof(int), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    *nwork = static_cast<uint64_t>(nw);
    return RL_OK;
}

int find_targets(rl_bh_ctx* c, uint32_t step) {
    if (!c->prepared) {
        return fail(c, RL_ESTATE, "find_target_neurons called before prepare_update_connectivity");
    }
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    c->nreq = 0;
    c->searched_valid = false;
    if (n == 0) {
        c->searched_valid = true;
        return RL_OK;
    }
    // slot offsets: exclusive prefix sum of the vacant axons
    k_widen<<<grid_for(n + 1), kBlock, 0, s>>>(n, c->vac[0], c->wide);
    RL_HIP(c, hipGetLastError());
    size_t tb = 0;
    RL_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, c->wide, c->slot_off, static_cast<int>(n) + 1, s));
    RL_TRY(ensure_tmp(c, tb));
    RL_HIP(c, hipcub::DeviceScan::ExclusiveSum(c->tmp, tb, c->wide, c->slot_off, static_cast<int>(n) + 1, s));
    uint64_t slots = 0;
    RL_HIP(c, hipMemcpyAsync(&slots, c->slot_off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    if (slots >= (1ULL << 32)) {
        return fail(c, RL_ELIMIT, "more than 2^32 - 1 vacant axons in one update");
    }
    if (slots > c->slots_cap) {
        RL_TRY(dalloc(c, &c->slot_target, slots));
        RL_TRY(dalloc(c, &c->slot_tests, slots));
        RL_TRY(dalloc(c, &c->slow_items, slots));
        c->slots_cap = slots;
    }
    c->slots = slots;
    unsigned long long* counters = c->counters;
    RL_HIP(c, hipMemsetAsync(counters, 0, 4 * sizeof(unsigned long long), s));
    if (slots > 0) {
        RL_HIP(c, hipMemsetAsync(c->slot_target, 0xFF, slots * sizeof(uint32_t), s));
        RL_HIP(c, hipMemsetAsync(c->slot_tests, 0, slots * sizeof(uint32_t), s));
    }

    SearchArgs a{};
    a.n = n;
    a.m = c->tree.m;
    a.order = c->order;
    a.pos = c->pos;
    a.qpos = c->qpos;
    a.signal = c->signal;
    a.enabled = c->enabled;
    a.vac_axon = c->vac[0];
    a.slot_off = c->slot_off;
    for (int k = 0; k < 2; k++) {
        a.rec64[k] = c->tree.rec64[k];
        a.rec32[k] = c->tree.rec32[k];
        a.meta[k] = c->tree.meta[k];
    }
    a.theta = c->cfg.theta;
    a.scale64 = -1.0 / (c->cfg.sigma * c->cfg.sigma);
    a.mu = c->cfg.mu;
    a.theta2f = static_cast<float>(c->cfg.theta * c->cfg.theta);
    a.k2f = static_cast<float>(-1.4426950408889634 / (c->cfg.sigma * c->cfg.sigma));
    a.muf = static_cast<float>(c->cfg.mu);
    a.net_of = c->net_of;
    a.net_off = c->net_off;
    a.net_seed = c->net_seed;
    a.net_k2q = c->net_k2q;
    a.net_invq = c->net_invq;
    a.guard = 2.0e-5f;
    a.seed = c->cfg.seed;
    a.step = step;
    a.slot_target = c->slot_target;
    a.slot_tests = c->slot_tests;
    a.found = c->found;
    a.searched = c->searched;
    a.counters = counters;
    const size_t lds = static_cast<size_t>(std::max<uint32_t>(c->tree.depth, 1U)) * kBlock * sizeof(uint32_t);
    RL_HIP(c, hipMemsetAsync(c->found, 0, n * sizeof(uint32_t), s));
    RL_HIP(c, hipMemsetAsync(c->searched, 0, n * sizeof(uint32_t), s));
    RL_HIP(c, hipEventRecord(c->ev0, s));
    const bool exact = (c->cfg.flags & RL_FLAG_EXACT_ORDER) != 0 || c->tree.depth == 0;
    const bool mixed = c->cfg.precision == RL_PREC_FP32_MIXED;
    uint64_t nlane = 0;
    a.redo = c->redo;
    a.n_redo = reinterpret_cast<unsigned int*>(c->coop_cursor + 9);
    RL_HIP(c, hipMemsetAsync(c->coop_cursor, 0, 10 * sizeof(unsigned long long), s));
    if (exact) {
        // per-lane kernel over the active sources in Morton order, then the fallback launch
        RL_TRY(compact_active(c, &nlane));
        a.order = c->coop_work;
        a.nwork = nlane;
        c->last_kernel_ms = 0.0;
        if (nlane > 0) {
            RL_HIP(c, hipEventRecord(c->ev2, s));
            if (mixed) {
                k_search<true, false><<<grid_for(nlane), kBlock, lds, s>>>(a);
            } else {
                k_search<false, false><<<grid_for(nlane), kBlock, lds, s>>>(a);
            }
            RL_HIP(c, hipGetLastError());
            RL_HIP(c, hipEventRecord(c->ev3, s));
        }
        unsigned int nredo = 0;
        RL_HIP(c, hipMemcpyAsync(&nredo, a.n_redo, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipStreamSynchronize(s));
        if (nlane > 0) {
            float kms = 0.0f;
            RL_HIP(c, hipEventElapsedTime(&kms, c->ev2, c->ev3));
            c->last_kernel_ms = kms;
        }
        nlane = nredo;
        a.order = c->redo;
    } else {
        RL_TRY(launch_search_coop(c, a, &nlane));
        a.order = c->coop_overflow;
    }
    a.nwork = nlane;
    if (nlane > 0) {
        if (mixed) {
            k_search<true, true><<<grid_for(nlane), kBlock, lds, s>>>(a);
        } else {
            k_search<false, true><<<grid_for(nlane), kBlock, lds, s>>>(a);
        }
        RL_HIP(c, hipGetLastError());
    }
    c->last_lane_sources = nlane;
    RL_HIP(c, hipEventRecord(c->ev1, s));
    // dense requests in (source, j) order
    k_widen<<<grid_for(n + 1), kBlock, 0, s>>>(n, c->found, c->wide);
    RL_HIP(c, hipGetLastError());
    RL_HIP(c, hipcub::DeviceScan::ExclusiveSum(c->tmp, tb, c->wide, c->req_off, static_cast<int>(n) + 1, s));
    uint64_t nreq = 0;
    unsigned long long hc[4] = { 0, 0, 0, 0 };
    RL_HIP(c, hipMemcpyAsync(&nreq, c->req_off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    float ms = 0.0f;
    RL_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_search_ms = ms;
    if (nreq > c->req_cap) {
        RL_TRY(dalloc(c, &c->req_target, nreq));
        RL_TRY(dalloc(c, &c->req_source, nreq));
        RL_TRY(dalloc(c, &c->req_signal, nreq));
        RL_TRY(dalloc(c, &c->req_j, nreq));
        RL_TRY(dalloc(c, &c->req_ok, nreq));
        c->req_cap = nreq;
    }
    c->nreq = nreq;
    if (nreq > 0) {
        k_emit_requests<<<grid_for(n), kBlock, 0, s>>>(n, c->found, c->req_off, c->slot_off, c->slot_target, c->signal,
                                                        c->req_target, c->req_source, c->req_signal, c->req_j);
        RL_HIP(c, hipGetLastError());
    }
    RL_HIP(c, hipStreamSynchronize(s));
    c->last_tests = hc[0];
    c->last_tests_ref = hc[1];
    c->last_searches = hc[2];
    c->searched_valid = true;
    return RL_OK;
}

}  // namespace rl
