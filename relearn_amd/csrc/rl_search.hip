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
#include <algorithm>

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

// Counter increments of a lane, added to the counter block once per thread block at the end of a
// kernel: an atomic per source on the same counters serialises.
struct Totals {
    unsigned long long exec = 0;
    unsigned long long ref = 0;
    unsigned long long searched = 0;
};

__device__ __forceinline__ unsigned long long wave_total(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    return v;
}

// Sums the lanes' totals over the block (all threads must call it) and adds them to the counters.
__device__ __forceinline__ void add_totals(unsigned long long* counters, const Totals& t) {
    __shared__ unsigned long long part[3][kBlock / 64];
    const unsigned long long v[3] = { wave_total(t.exec), wave_total(t.ref), wave_total(t.searched) };
    const uint32_t wv = threadIdx.x / 64U;
    if ((threadIdx.x & 63U) == 0) {
        for (int k = 0; k < 3; k++) {
            part[k][wv] = v[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s[3] = { 0, 0, 0 };
        for (uint32_t w = 0; w < blockDim.x / 64U; w++) {
            for (int k = 0; k < 3; k++) {
                s[k] += part[k][w];
            }
        }
        if (s[0] != 0) atomicAdd(&counters[kCtrTestsExec], s[0]);
        if (s[1] != 0) atomicAdd(&counters[kCtrTestsRef], s[1]);
        if (s[2] != 0) atomicAdd(&counters[kCtrSearches], s[2]);
    }
}

// The whole search of source src by one lane; stk is the lane's LDS stack column (stride apart).
template <bool kMixed, bool kFallback>
__device__ __forceinline__ void search_source(const SearchArgs& a, uint32_t src, uint32_t* stk, uint32_t stride,
                                              Totals& tot) {
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
        tot.searched += f == na ? na : 1;
        return;
    }
    // level 1: one list for all axons of the neuron
    const ListSum l1 = list_sum<kMixed, kFallback>(a, L, L.root, stk, stride);
    if (l1.redo) {
        a.redo[atomicAdd(a.n_redo, 1ULL)] = src;
        return;
    }
    exec += l1.tests_exec;
    if (l1.count == 0 || l1.sum == 0.0) {
        a.slot_tests[slot0] = l1.tests_ref;
        a.found[src] = 0;
        a.searched[src] = 1;
        tot.exec += exec;
        tot.ref += l1.tests_ref;
        tot.searched += 1;
        return;
    }
    const double next1 = nextafter(l1.sum, l1.sum + 0.00001);
    for (uint32_t j0 = 0; j0 < na; j0 += kDrawsPerPass) {
        PickVisitor pv;
        pv.nd = static_cast<int>(min(na - j0, static_cast<uint32_t>(kDrawsPerPass)));
#pragma unroll
        for (int k = 0; k < kDrawsPerPass; k++) {
            pv.pick[k] = kNone;
            pv.r[k] = k < pv.nd ? draw_to_uniform(pick_draw(L.seed, L.local, L.step, j0 + k, 1U), next1) : 0.0;
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
                a.redo[atomicAdd(a.n_redo, 1ULL)] = src;
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
            pv.r[0] = draw_to_uniform(pick_draw(L.seed, L.local, L.step, j, level), nextafter(ls.sum, ls.sum + 0.00001));
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
    tot.exec += exec;
    tot.ref += ref;
    tot.searched += searched;
}

// Sources a.order[0 .. *a.nwork), grid-stride, one lane per source.
template <bool kMixed, bool kFallback>
__global__ void __launch_bounds__(kBlock) k_search(SearchArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t count = *a.nwork;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    Totals tot;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < count; t += stride) {
        search_source<kMixed, kFallback>(a, a.order[t], lds_stack + threadIdx.x, blockDim.x, tot);
    }
    add_totals(a.counters, tot);
}

// Descents that the cooperative kernels hand over (a level that needs expansion or the reference's
// fallbacks): one lane continues the search of one slot from (node, level) with the general
// traversal, as the descent loop of k_search does.
template <bool kMixed>
__device__ __forceinline__ void descend_item(const SearchArgs& a, const uint4 it, uint32_t* stk, uint32_t stride,
                                             Totals& tot) {
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
        pv.r[0] = draw_to_uniform(pick_draw(L.seed, L.local, L.step, j, level), nextafter(ls.sum, ls.sum + 0.00001));
        exec += list_pick<kMixed, true>(a, L, node, ls, pv, stk, stride);
        node = pv.pick[0];
    }
    a.slot_tests[slot] += tests;
    a.slot_target[slot] = ok ? node : kNone;
    tot.exec += exec;
}

template <bool kMixed>
__global__ void __launch_bounds__(kBlock) k_descend_lane(SearchArgs a, const uint4* __restrict__ items,
                                                         const unsigned long long* __restrict__ n_items) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t count = *n_items;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    Totals tot;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < count; t += stride) {
        descend_item<kMixed>(a, items[t], lds_stack + threadIdx.x, blockDim.x, tot);
    }
    add_totals(a.counters, tot);
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

// Start of a search: results and counters cleared, slots of this update initialised.  When the
// update has more vacant axons than the slot buffers hold, kCtrOverCap is raised and no source is
// searched (the update then changes nothing).
__global__ void k_search_reset(uint64_t n, const uint64_t* __restrict__ slot_total, uint64_t cap,
                               uint32_t* __restrict__ found, uint32_t* __restrict__ searched,
                               uint32_t* __restrict__ slot_target, uint32_t* __restrict__ slot_tests,
                               unsigned long long* __restrict__ counters, unsigned long long* __restrict__ cursor) {
    const uint64_t total = *slot_total;
    const bool over = total > cap;
    const uint64_t slots = over ? 0 : total;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t t0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    for (uint64_t i = t0; i < n; i += stride) {
        found[i] = 0;
        searched[i] = 0;
    }
    for (uint64_t i = t0; i < slots; i += stride) {
        slot_target[i] = kNone;
        slot_tests[i] = 0;
    }
    if (t0 < kCounters) {
        if (t0 == kCtrSticky) {
            if (over) {
                counters[t0] += 1;
            }
        } else if (t0 != kCtrDeleted && t0 != kCtrDelShort) {  // those belong to the element commit before
            counters[t0] = (t0 == kCtrOverCap && over) ? 1ULL : 0ULL;
        }
    }
    if (t0 < kCursorSlots) {
        cursor[t0] = 0;
    }
}

__global__ void k_active_flags(uint64_t n, const uint32_t* __restrict__ order, const uint8_t* __restrict__ enabled,
                               const uint32_t* __restrict__ vac, const unsigned long long* __restrict__ counters,
                               uint32_t* __restrict__ flag) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t < n) {
        const uint32_t src = order[t];
        flag[t] = (counters[kCtrOverCap] == 0 && enabled[src] != 0 && vac[src] != 0) ? 1U : 0U;
    }
}

__global__ void k_scatter_active(uint64_t n, const uint32_t* __restrict__ flag, const uint64_t* __restrict__ off,
                                 const uint32_t* __restrict__ order, uint32_t* __restrict__ work) {
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t < n && flag[t] != 0) {
        work[off[t]] = order[t];
    }
}

// Writes the accepted nodes of a traversal with their probabilities, up to a capacity; counts all.
struct RecordVisitor {
    uint32_t* nodes;
    double* probs;
    uint64_t cap;
    uint64_t count = 0;
    __device__ __forceinline__ void operator()(uint32_t node, double p) {
        if (count < cap) {
            nodes[count] = node;
            probs[count] = p;
        }
        count++;
    }
};

// Level-1 candidate list of each listed source (one lane per source): the list the search sums and
// picks from, with the probabilities of the context's precision.
template <bool kMixed>
__global__ void __launch_bounds__(kBlock) k_candidates(SearchArgs a, const uint32_t* __restrict__ sources,
                                                       uint64_t count, uint64_t cap, uint32_t* __restrict__ nodes,
                                                       double* __restrict__ probs, uint32_t* __restrict__ counts) {
    extern __shared__ uint32_t lds_stack[];
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (t >= count) {
        return;
    }
    Lane L;
    lane_source(a, L, sources[t], kMixed);
    RecordVisitor v;
    v.nodes = nodes + t * cap;
    v.probs = probs + t * cap;
    v.cap = cap;
    (void)traverse<kMixed>(a, L, L.root, kGauss, lds_stack + threadIdx.x, blockDim.x, v);
    counts[t] = static_cast<uint32_t>(v.count);
}

__global__ void k_collect(unsigned long long* __restrict__ counters, const uint64_t* __restrict__ slots,
                          const uint64_t* __restrict__ requests, const uint64_t* __restrict__ active) {
    if (threadIdx.x == 0) {
        counters[kCtrSlots] = *slots;
        counters[kCtrRequests] = *requests;
        counters[kCtrActive] = *active;
    }
}

}  // namespace

int launch_descend_lane(rl_bh_ctx* c, const SearchArgs& sa, const uint4* items, const unsigned long long* count) {
    hipStream_t s = c->stream;
    const size_t lds = static_cast<size_t>(std::max<uint32_t>(c->tree.depth, 1U)) * kBlock * sizeof(uint32_t);
    if (c->cfg.precision == RL_PREC_FP32_MIXED) {
        k_descend_lane<true><<<kStreamBlocks / 2, kBlock, lds, s>>>(sa, items, count);
    } else {
        k_descend_lane<false><<<kStreamBlocks / 2, kBlock, lds, s>>>(sa, items, count);
    }
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

// Slot and request buffers for `want` vacant axons.  Grows with headroom, so that updates whose
// vacancies vary do not reallocate every time; a captured period refers to the old buffers and is
// released.
int ensure_slots(rl_bh_ctx* c, uint64_t want) {
    if (want <= c->slots_cap && c->slot_target != nullptr) {
        return RL_OK;
    }
    if (c->capturing) {
        return fail(c, RL_ESTATE, "slot buffers cannot grow during graph capture");
    }
    constexpr uint64_t kMaxSlots = 0xFFFFFFFFULL;
    if (want > kMaxSlots) {
        return fail(c, RL_ELIMIT, "more than 2^32 - 1 vacant axons in one update");
    }
    const uint64_t cap = std::min<uint64_t>(kMaxSlots, want + want / 8 + 1024);
    destroy_graph(c);
    RL_HIP(c, hipStreamSynchronize(c->stream));
    RL_TRY(dalloc(c, &c->slot_target, cap));
    RL_TRY(dalloc(c, &c->slot_tests, cap));
    RL_TRY(dalloc(c, &c->slow_items, cap));
    RL_TRY(dalloc(c, &c->req_target, cap));
    RL_TRY(dalloc(c, &c->req_source, cap));
    RL_TRY(dalloc(c, &c->req_signal, cap));
    RL_TRY(dalloc(c, &c->req_j, cap));
    RL_TRY(dalloc(c, &c->req_ok, cap));
    RL_TRY(dalloc(c, &c->c_prio, cap));
    RL_TRY(dalloc(c, &c->c_order, cap));
    RL_TRY(dalloc(c, &c->c_sprio, cap));
    RL_TRY(dalloc(c, &c->c_sidx, cap));
    c->slots_cap = cap;
    return RL_OK;
}

// The search arguments of the context's current tree, buffers and parameters.
SearchArgs search_args(rl_bh_ctx* c) {
    SearchArgs a{};
    a.n = c->n;
    a.m = c->tree.m;
    a.pos = c->pos;
    a.qpos = c->qpos;
    a.signal = c->signal;
    a.enabled = c->enabled;
    a.vac_axon = c->vac[0];
    a.slot_off = c->slot_off;
    for (int k = 0; k < 2; k++) {
        a.rec64[k] = c->tree.rec64[k];
        a.rec32[k] = c->tree.rec32[k];
        a.wrec[k] = c->tree.wrec[k];
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
    {
        const double t2 = c->cfg.theta * c->cfg.theta;
        a.guard_abs = static_cast<float>(12.0 * t2 / 2.0e-5 + 6.0 * t2);
    }
    a.seed = c->cfg.seed;
    a.step = c->d_steps + 1;
    a.slot_target = c->slot_target;
    a.slot_tests = c->slot_tests;
    a.found = c->found;
    a.searched = c->searched;
    a.counters = c->counters;
    a.redo = c->redo;
    a.n_redo = c->counters + kCtrRedo;
    return a;
}

// Sizes the slot buffers for the resident vacancies (one host read of their sum).
int search_capacity(rl_bh_ctx* c) {
    RL_TRY(scan_counts(c, c->vac[0], c->n, c->slot_off));
    uint64_t total = 0;
    RL_HIP(c, hipMemcpyAsync(&total, c->slot_off + c->n, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    return ensure_slots(c, total);
}

// The search of one update on the context's stream, driven by counts in device memory: slot
// offsets, the active sources in Morton order, the search kernels, the dense requests.  No host
// synchronisation, allocation or memset, so it can be captured into a graph.  The update index is
// read from d_steps[1].
int enqueue_search(rl_bh_ctx* c) {
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    if (n == 0) {
        return RL_OK;
    }
    RL_TRY(scan_counts(c, c->vac[0], n, c->slot_off));
    k_search_reset<<<kStreamBlocks, kBlock, 0, s>>>(n, c->slot_off + n, c->slots_cap, c->found, c->searched,
                                                    c->slot_target, c->slot_tests, c->counters, c->coop_cursor);
    RL_HIP(c, hipGetLastError());
    const bool exact = (c->cfg.flags & RL_FLAG_EXACT_ORDER) != 0 || c->tree.depth == 0;
    // the group walk builds its own source lists (per dendrite type); the other kernels take the
    // active sources in Morton order
    const bool walk = !exact && uses_walk(c);
    const uint64_t* n_active = c->wide + n;
    if (walk) {
        n_active = c->walk_off + 2 * n;
    } else {
        k_active_flags<<<grid_for(n), kBlock, 0, s>>>(n, c->order, c->enabled, c->vac[0], c->counters,
                                                       c->coop_flag);
        RL_HIP(c, hipGetLastError());
        RL_TRY(scan_counts(c, c->coop_flag, n, c->wide));
        k_scatter_active<<<grid_for(n), kBlock, 0, s>>>(n, c->coop_flag, c->wide, c->order, c->coop_work);
        RL_HIP(c, hipGetLastError());
    }

    SearchArgs a = search_args(c);
    a.order = c->coop_work;
    a.nwork = c->wide + n;
    const size_t lds = static_cast<size_t>(std::max<uint32_t>(c->tree.depth, 1U)) * kBlock * sizeof(uint32_t);
    const bool mixed = c->cfg.precision == RL_PREC_FP32_MIXED;
    const bool timed = !c->capturing;
    if (exact) {
        // one lane per active source in Morton order, then the fallback pass over the redo list
        if (timed) RL_HIP(c, hipEventRecord(c->ev2, s));
        if (mixed) {
            k_search<true, false><<<grid_for(n), kBlock, lds, s>>>(a);
        } else {
            k_search<false, false><<<grid_for(n), kBlock, lds, s>>>(a);
        }
        RL_HIP(c, hipGetLastError());
        if (timed) RL_HIP(c, hipEventRecord(c->ev3, s));
        a.order = c->redo;
        a.nwork = reinterpret_cast<const uint64_t*>(c->counters + kCtrRedo);
    } else {
        if (walk) {
            RL_TRY(launch_search_walk(c, a));
        } else {
            RL_TRY(launch_search_coop(c, a));
        }
        a.order = c->coop_overflow;
        a.nwork = reinterpret_cast<const uint64_t*>(c->counters + kCtrOverflow);
    }
    if (mixed) {
        k_search<true, true><<<kStreamBlocks / 8, kBlock, lds, s>>>(a);
    } else {
        k_search<false, true><<<kStreamBlocks / 8, kBlock, lds, s>>>(a);
    }
    RL_HIP(c, hipGetLastError());
    // dense requests in (source, j) order
    RL_TRY(scan_counts(c, c->found, n, c->req_off));
    k_emit_requests<<<grid_for(n), kBlock, 0, s>>>(n, c->found, c->req_off, c->slot_off, c->slot_target, c->signal,
                                                    c->req_target, c->req_source, c->req_signal, c->req_j);
    RL_HIP(c, hipGetLastError());
    k_collect<<<1, 64, 0, s>>>(c->counters, c->slot_off + n, c->req_off + n, n_active);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

int read_counters(rl_bh_ctx* c, unsigned long long* h) {
    RL_HIP(c, hipMemcpyAsync(h, c->counters, kCounters * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    if (h[kCtrSticky] != 0) {
        RL_TRY(fill_u32(c, reinterpret_cast<uint32_t*>(c->counters + kCtrSticky), 0U, 2, nullptr, 0));
    }
    return RL_OK;
}

int find_targets(rl_bh_ctx* c, uint32_t step) {
    if (!c->prepared) {
        return fail(c, RL_ESTATE, "find_target_neurons called before prepare_update_connectivity");
    }
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    c->nreq = 0;
    c->slots = 0;
    c->searched_valid = false;
    c->committed = false;
    if (n == 0) {
        c->searched_valid = true;
        return RL_OK;
    }
    RL_TRY(search_capacity(c));
    RL_TRY(set_steps(c, 0, step));
    RL_HIP(c, hipEventRecord(c->ev0, s));
    RL_TRY(enqueue_search(c));
    RL_HIP(c, hipEventRecord(c->ev1, s));
    unsigned long long h[kCounters];
    RL_TRY(read_counters(c, h));
    if (h[kCtrOverCap] != 0) {
        return fail(c, RL_ELIMIT, "vacant axons exceed the slot buffers");
    }
    float ms = 0.0f, kms = 0.0f;
    RL_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    RL_HIP(c, hipEventElapsedTime(&kms, c->ev2, c->ev3));
    c->last_search_ms = ms;
    c->last_kernel_ms = h[kCtrActive] > 0 ? kms : 0.0;
    c->slots = h[kCtrSlots];
    c->nreq = h[kCtrRequests];
    c->last_tests = h[kCtrTestsExec];
    c->last_tests_ref = h[kCtrTestsRef];
    c->last_searches = h[kCtrSearches];
    c->last_active = h[kCtrActive];
    c->last_slow_descents = h[kCtrSlow];
    c->last_lane_sources = (c->cfg.flags & RL_FLAG_EXACT_ORDER) != 0 || c->tree.depth == 0 ? h[kCtrRedo] : h[kCtrOverflow];
    c->last_walk_steps = h[kCtrWalkSteps];
    c->last_walk_slots = h[kCtrWalkSlots];
    c->last_walk_bytes = h[kCtrWalkBytes];
    c->searched_valid = true;
    return RL_OK;
}

int candidate_lists(rl_bh_ctx* c, uint64_t num_sources, const uint32_t* sources, uint64_t cap, uint32_t* nodes,
                    double* probs, uint32_t* counts) {
    if (!c->prepared) {
        return fail(c, RL_ESTATE, "rl_bh_candidate_lists: prepare the update first");
    }
    for (uint64_t i = 0; i < num_sources; i++) {
        if (sources[i] >= c->n) {
            return fail(c, RL_EINVAL, "rl_bh_candidate_lists: source out of range");
        }
    }
    if (num_sources == 0) {
        return RL_OK;
    }
    hipStream_t s = c->stream;
    uint32_t* d_src = nullptr;
    uint32_t* d_nodes = nullptr;
    double* d_probs = nullptr;
    uint32_t* d_counts = nullptr;
    const uint64_t total = num_sources * std::max<uint64_t>(cap, 1);
    int status = RL_OK;
    do {
        if ((status = dalloc(c, &d_src, num_sources)) != RL_OK) break;
        if ((status = dalloc(c, &d_nodes, total)) != RL_OK) break;
        if ((status = dalloc(c, &d_probs, total)) != RL_OK) break;
        if ((status = dalloc(c, &d_counts, num_sources)) != RL_OK) break;
        hipError_t e = hipMemcpyAsync(d_src, sources, num_sources * sizeof(uint32_t), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            status = hip_fail(c, e, "hipMemcpyAsync");
            break;
        }
        const SearchArgs a = search_args(c);
        const size_t lds = static_cast<size_t>(std::max<uint32_t>(c->tree.depth, 1U)) * kBlock * sizeof(uint32_t);
        if (c->cfg.precision == RL_PREC_FP32_MIXED) {
            k_candidates<true><<<grid_for(num_sources), kBlock, lds, s>>>(a, d_src, num_sources, cap, d_nodes, d_probs,
                                                                           d_counts);
        } else {
            k_candidates<false><<<grid_for(num_sources), kBlock, lds, s>>>(a, d_src, num_sources, cap, d_nodes, d_probs,
                                                                            d_counts);
        }
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = hipMemcpyAsync(counts, d_counts, num_sources * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
            (cap > 0 && (e = hipMemcpyAsync(nodes, d_nodes, num_sources * cap * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) != hipSuccess) ||
            (cap > 0 && (e = hipMemcpyAsync(probs, d_probs, num_sources * cap * sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) ||
            (e = hipStreamSynchronize(s)) != hipSuccess) {
            status = hip_fail(c, e, "candidate lists");
        }
    } while (false);
    (void)hipStreamSynchronize(s);
    dfree(&d_src);
    dfree(&d_nodes);
    dfree(&d_probs);
    dfree(&d_counts);
    return status;
}

}  // namespace rl
