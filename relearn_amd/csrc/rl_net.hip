// Network graph of the synaptic input on the device: creation and deletion of synapses.
//
// The synaptic input of a neuron (SynapticEquallyWeightedActivityInputCPU::update_local_input,
// source/neurons/input/SynapticEquallyWeightedActivityInputCPU.cpp:21-50) sums the weights of its
// in-edges (NetworkGraph::get_local_in_edges) whose source fired in the last step.  The device keeps
// those in-edges as a CSR by target with one entry per (source, target) pair whose weight is the sum
// of the pair's synapses, as NetworkGraph::add_synapse merges them; a pair whose weight returns to 0
// is dropped at the next append.
//
// Creation: every commit appends the synapses it created (Neurons::update_connectivity adds them to
// the graph, source/neurons/Neurons.cpp:458-476), weight +1 for an excitatory source and -1 for an
// inhibitory one (ForwardConnector::process_requests, source/algorithm/Connector.cpp:33-121).  The
// new synapses are counted per target, prefix-summed and staged; one lane per target then merges
// them into the target's pairs, writing the merged list into the second buffer, and the compacted
// graph is copied back.  Counts come from device memory and nothing allocates, so the append can be
// captured with the commit; an append that would not fit appends nothing and raises a device flag.
//
// Deletion (SynapseDeletionFinderCPU::delete_synapses, source/neurons/helper/SynapseDeletionFinderCPU.cpp,
// with RandomSynapseDeletionFinder::find_synapses_on_neuron, SynapseDeletionFinder.cpp:53-69): after
// the element commit of one type, a neuron that lost k bound elements of that type deletes k of its
// synapse units of that kind, drawn uniformly without replacement: axons its outgoing synapses,
// dendrites its incoming synapses of the dendrite's sign; a pair of weight w counts |w| units
// (SynapseDeletionFinderBase::register_synapses, SynapseDeletionFinderBase.cpp:33-151).  Every deleted
// unit lowers the pair's weight by one and disconnects one element of the partner
// (commit_deletions, ElementBaseBase::disconnect_elements, ElementBaseBase.h:311-319).  The draw is
// counter-based: the k units with the smallest deletion_key(...) over (partner, copy), so the subset
// does not depend on the order the pairs are stored in.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rl_device.h"
#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

__device__ __forceinline__ int32_t unit_sign(int32_t w) {
    return w < 0 ? -1 : 1;
}

// ---------------------------------------------------------------- append

__global__ void k_net_count(const uint64_t* __restrict__ n_req, const uint8_t* __restrict__ ok,
                            const uint32_t* __restrict__ tgt, uint32_t* __restrict__ ncnt) {
    const uint64_t m = *n_req;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        if (ok[i] != 0) {
            atomicAdd(&ncnt[tgt[i]], 1U);
        }
    }
}

// The merge writes target t's list at off[t] + npre[t] of the second buffer, room for its old entries
// and its new synapses; that layout must fit the capacity.
__global__ void k_net_check(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ npre,
                            uint64_t cap, unsigned long long* __restrict__ counters) {
    if (threadIdx.x == 0) {
        const uint64_t add = npre[n];
        const bool over = off[n] + add > cap;
        counters[kCtrNetOver] = over ? 1ULL : 0ULL;
        counters[kCtrNetNew] = over ? 0ULL : add;
        if (over) {
            counters[kCtrSticky] += 1ULL;
        }
    }
}

// New synapses staged per target at npre[t]: their sources (the weight follows from the source's
// signal type).  ncnt returns to zero.
__global__ void k_net_stage(const uint64_t* __restrict__ n_req, const uint8_t* __restrict__ ok,
                            const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ rsrc,
                            const uint64_t* __restrict__ npre, uint32_t* __restrict__ ncnt, uint32_t* __restrict__ stage,
                            const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t m = *n_req;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        if (ok[i] == 0) {
            continue;
        }
        const uint32_t t = tgt[i];
        stage[npre[t] + atomicSub(&ncnt[t], 1U) - 1U] = rsrc[i];
    }
}

// One lane per target: its non-zero pairs, then its new synapses added to the pair of their source
// or appended as a new pair.  ncnt[t] gets the merged count.
__global__ void k_net_merge(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ npre,
                            const uint32_t* __restrict__ src, const int32_t* __restrict__ w,
                            const uint32_t* __restrict__ stage, const uint8_t* __restrict__ signal,
                            uint32_t* __restrict__ src2, int32_t* __restrict__ w2, uint32_t* __restrict__ ncnt,
                            const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        const uint64_t base = off[t] + npre[t];
        uint32_t k = 0;
        for (uint64_t e = off[t]; e < off[t + 1]; e++) {
            if (w[e] != 0) {
                src2[base + k] = src[e];
                w2[base + k] = w[e];
                k++;
            }
        }
        for (uint64_t q = npre[t]; q < npre[t + 1]; q++) {
            const uint32_t s = stage[q];
            const int32_t dw = signal[s] != 0 ? -1 : 1;
            uint32_t j = 0;
            while (j < k && src2[base + j] != s) {
                j++;
            }
            if (j < k) {
                w2[base + j] += dw;
            } else {
                src2[base + k] = s;
                w2[base + k] = dw;
                k++;
            }
        }
        ncnt[t] = k;
    }
}

// Compacted lists back into the primary arrays, at the prefix sum noff of the merged counts.
__global__ void k_net_copyback(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ npre,
                               const uint64_t* __restrict__ noff, const uint32_t* __restrict__ src2,
                               const int32_t* __restrict__ w2, uint32_t* __restrict__ src, int32_t* __restrict__ w,
                               const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        const uint64_t from = off[t] + npre[t];
        const uint64_t to = noff[t];
        const uint64_t k = noff[t + 1] - to;
        for (uint64_t j = 0; j < k; j++) {
            src[to + j] = src2[from + j];
            w[to + j] = w2[from + j];
        }
    }
}

__global__ void k_net_offsets(uint64_t n, const uint64_t* __restrict__ noff, uint64_t* __restrict__ off,
                              const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t <= n; t += stride) {
        off[t] = noff[t];
    }
}

// ---------------------------------------------------------------- synaptic input by push

// Non-zero pairs per source (out_cnt starts at zero).
__global__ void k_out_count(uint64_t n, const uint64_t* __restrict__ off, const uint32_t* __restrict__ src,
                            const int32_t* __restrict__ w, uint32_t* __restrict__ out_cnt) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        for (uint64_t e = off[t]; e < off[t + 1]; e++) {
            if (w[e] != 0) {
                atomicAdd(&out_cnt[src[e]], 1U);
            }
        }
    }
}

// The pairs grouped by source at out_off (order inside a source arbitrary); out_cnt returns to zero.
__global__ void k_out_scatter(uint64_t n, const uint64_t* __restrict__ off, const uint32_t* __restrict__ src,
                              const int32_t* __restrict__ w, const uint64_t* __restrict__ out_off,
                              uint32_t* __restrict__ out_cnt, uint32_t* __restrict__ out_tgt, int32_t* __restrict__ out_w) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        for (uint64_t e = off[t]; e < off[t + 1]; e++) {
            const int32_t we = w[e];
            if (we != 0) {
                const uint32_t s = src[e];
                const uint64_t pos = out_off[s] + atomicSub(&out_cnt[s], 1U) - 1U;
                out_tgt[pos] = static_cast<uint32_t>(t);
                out_w[pos] = we;
            }
        }
    }
}

// acc[t] += the weights of the pairs whose source has fired set (acc starts at zero).
__global__ void k_syn_push_all(uint64_t n, const uint8_t* __restrict__ fired, const uint64_t* __restrict__ out_off,
                               const uint32_t* __restrict__ out_tgt, const int32_t* __restrict__ out_w,
                               int32_t* __restrict__ acc) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < n; s += stride) {
        if (fired[s] != 0) {
            for (uint64_t e = out_off[s]; e < out_off[s + 1]; e++) {
                atomicAdd(&acc[out_tgt[e]], out_w[e]);
            }
        }
    }
}

// ---------------------------------------------------------------- deletion

struct DelArgs {
    uint64_t n;
    const uint32_t* net_of;
    const uint64_t* net_off;
    const uint64_t* net_seed;
    const uint32_t* step;        // device: plasticity update index
    const uint32_t* deleted;     // units to delete per neuron (this type)
    const uint8_t* signal;
    const uint64_t* off;         // graph
    const uint32_t* src;
    int32_t* w;
    uint32_t* conn[3];
    uint32_t* vac[3];
    unsigned long long* counters;
};

// A unit (partner, copy) of a neuron with its key; units compare by (key, partner, copy).
struct Unit {
    uint64_t key;
    uint32_t partner;
    uint32_t copy;
};

__device__ __forceinline__ bool unit_less(const Unit& a, const Unit& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.partner != b.partner) return a.partner < b.partner;
    return a.copy < b.copy;
}

// Lists of the out-entries of the sources that delete axons: counted, prefix-summed, scattered as
// (entry, target).
__global__ void k_del_out_count(uint64_t n, const uint64_t* __restrict__ off, const uint32_t* __restrict__ src,
                                const int32_t* __restrict__ w, const uint32_t* __restrict__ deleted,
                                uint32_t* __restrict__ ocnt) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        for (uint64_t e = off[t]; e < off[t + 1]; e++) {
            const uint32_t s = src[e];
            if (w[e] != 0 && deleted[s] != 0) {
                atomicAdd(&ocnt[s], 1U);
            }
        }
    }
}

__global__ void k_del_out_scatter(uint64_t n, const uint64_t* __restrict__ off, const uint32_t* __restrict__ src,
                                  const int32_t* __restrict__ w, const uint32_t* __restrict__ deleted,
                                  const uint64_t* __restrict__ ooff, uint32_t* __restrict__ ocnt,
                                  uint2* __restrict__ olist) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < n; t += stride) {
        for (uint64_t e = off[t]; e < off[t + 1]; e++) {
            const uint32_t s = src[e];
            if (w[e] != 0 && deleted[s] != 0) {
                olist[ooff[s] + atomicSub(&ocnt[s], 1U) - 1U] = make_uint2(static_cast<uint32_t>(e), static_cast<uint32_t>(t));
            }
        }
    }
}

// The k-th smallest unit of a neuron (k >= 1), by k selection passes over its entries.  `entry(j)`
// gives (partner, |weight|) of entry j < m, or |weight| 0 for entries that do not count.
template <typename Entry>
__device__ Unit kth_unit(uint64_t seed, uint32_t local, uint32_t step, uint32_t type, uint32_t m, uint32_t k,
                         Entry entry) {
    Unit prev{ 0, 0, 0 };
    bool have_prev = false;
    for (uint32_t pass = 0; pass < k; pass++) {
        Unit best{ ~0ULL, ~0U, ~0U };
        bool found = false;
        for (uint32_t j = 0; j < m; j++) {
            const uint2 pw = entry(j);
            for (uint32_t c = 0; c < pw.y; c++) {
                const Unit u{ deletion_key(seed, local, step, pw.x, type, c), pw.x, c };
                if ((!have_prev || unit_less(prev, u)) && (!found || unit_less(u, best))) {
                    best = u;
                    found = true;
                }
            }
        }
        prev = best;
        have_prev = true;
    }
    return prev;
}

__device__ __forceinline__ uint32_t local_id(const DelArgs& a, uint32_t i) {
    return i - static_cast<uint32_t>(a.net_off[a.net_of[i]]);
}

// Axons: source s deletes deleted[s] units among its outgoing pairs; each unit lowers the pair's
// weight and disconnects one dendrite (of s's signal type) of the target.
__global__ void k_del_axons(DelArgs a, const uint64_t* __restrict__ ooff, const uint2* __restrict__ olist) {
    const uint32_t step = *a.step;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long done = 0, short_of = 0;
    for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < a.n; s += stride) {
        const uint32_t want = a.deleted[s];
        if (want == 0) {
            continue;
        }
        const uint64_t o = ooff[s];
        const uint32_t m = static_cast<uint32_t>(ooff[s + 1] - o);
        const uint32_t g = a.net_of[s];
        const uint64_t seed = a.net_seed[g];
        const uint32_t self = local_id(a, static_cast<uint32_t>(s));
        auto entry = [&](uint32_t j) -> uint2 {
            const uint2 et = olist[o + j];
            const int32_t wt = a.w[et.x];
            return make_uint2(local_id(a, et.y), static_cast<uint32_t>(wt < 0 ? -wt : wt));
        };
        uint32_t units = 0;
        for (uint32_t j = 0; j < m; j++) {
            units += entry(j).y;
        }
        const uint32_t k = want < units ? want : units;
        short_of += want - k;
        if (k == 0) {
            continue;
        }
        const Unit last = kth_unit(seed, self, step, 0U, m, k, entry);
        const int sig = a.signal[s] != 0 ? 1 : 0;
        for (uint32_t j = 0; j < m; j++) {
            const uint2 et = olist[o + j];
            const uint2 pw = entry(j);
            uint32_t sel = 0;
            for (uint32_t c = 0; c < pw.y; c++) {
                const Unit u{ deletion_key(seed, self, step, pw.x, 0U, c), pw.x, c };
                sel += unit_less(last, u) ? 0U : 1U;
            }
            if (sel != 0) {
                a.w[et.x] -= unit_sign(a.w[et.x]) * static_cast<int32_t>(sel);
                atomicSub(&a.conn[1 + sig][et.y], sel);
                atomicAdd(&a.vac[1 + sig][et.y], sel);
                done += sel;
            }
        }
    }
    if (done != 0) atomicAdd(&a.counters[kCtrDeleted], done);
    if (short_of != 0) atomicAdd(&a.counters[kCtrDelShort], short_of);
}

// Dendrites of `type` (1 excitatory: pairs of positive weight, 2 inhibitory: negative): target t
// deletes deleted[t] units among its incoming pairs of that sign; each disconnects one axon of the
// source.
__global__ void k_del_dendrites(DelArgs a, uint32_t type) {
    const uint32_t step = *a.step;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long done = 0, short_of = 0;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < a.n; t += stride) {
        const uint32_t want = a.deleted[t];
        if (want == 0) {
            continue;
        }
        const uint64_t o = a.off[t];
        const uint32_t m = static_cast<uint32_t>(a.off[t + 1] - o);
        const uint32_t g = a.net_of[t];
        const uint64_t seed = a.net_seed[g];
        const uint32_t self = local_id(a, static_cast<uint32_t>(t));
        auto entry = [&](uint32_t j) -> uint2 {
            const int32_t wt = a.w[o + j];
            const bool match = type == 1 ? wt > 0 : wt < 0;
            return make_uint2(local_id(a, a.src[o + j]), match ? static_cast<uint32_t>(wt < 0 ? -wt : wt) : 0U);
        };
        uint32_t units = 0;
        for (uint32_t j = 0; j < m; j++) {
            units += entry(j).y;
        }
        const uint32_t k = want < units ? want : units;
        short_of += want - k;
        if (k == 0) {
            continue;
        }
        const Unit last = kth_unit(seed, self, step, type, m, k, entry);
        for (uint32_t j = 0; j < m; j++) {
            const uint2 pw = entry(j);
            uint32_t sel = 0;
            for (uint32_t c = 0; c < pw.y; c++) {
                const Unit u{ deletion_key(seed, self, step, pw.x, type, c), pw.x, c };
                sel += unit_less(last, u) ? 0U : 1U;
            }
            if (sel != 0) {
                a.w[o + j] -= unit_sign(a.w[o + j]) * static_cast<int32_t>(sel);
                const uint32_t s = a.src[o + j];
                atomicSub(&a.conn[0][s], sel);
                atomicAdd(&a.vac[0][s], sel);
                done += sel;
            }
        }
    }
    if (done != 0) atomicAdd(&a.counters[kCtrDeleted], done);
    if (short_of != 0) atomicAdd(&a.counters[kCtrDelShort], short_of);
}

int net_alloc_edges(rl_bh_ctx* c, uint64_t cap) {
    NetGraph& g = c->net;
    RL_TRY(dalloc(c, &g.src, cap));
    RL_TRY(dalloc(c, &g.w, cap));
    RL_TRY(dalloc(c, &g.src2, cap));
    RL_TRY(dalloc(c, &g.w2, cap));
    RL_TRY(dalloc(c, &g.olist, cap));
    RL_TRY(dalloc(c, &g.out_tgt, cap));
    RL_TRY(dalloc(c, &g.out_w, cap));
    g.cap = cap;
    return RL_OK;
}

// Counts the accepted requests of the last commit per target into ncnt and their prefix into npre.
int enqueue_net_count(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    RL_TRY(fill_u32(c, g.ncnt, 0U, n, nullptr, 0));
    k_net_count<<<kStreamBlocks, kBlock, 0, c->stream>>>(c->req_off + n, c->req_ok, c->req_target, g.ncnt);
    RL_HIP(c, hipGetLastError());
    return scan_counts(c, g.ncnt, n, g.npre);
}

// Stage, merge, compact.  The staging area is the commit's grouped-index scratch (slots_cap entries,
// at least the accepted requests), free once the commit has run.
int enqueue_net_merge(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    k_net_check<<<1, 64, 0, s>>>(n, g.off, g.npre, g.cap, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_stage<<<kStreamBlocks, kBlock, 0, s>>>(c->req_off + n, c->req_ok, c->req_target, c->req_source, g.npre,
                                                 g.ncnt, c->c_sidx, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_merge<<<grid_for(n), kBlock, 0, s>>>(n, g.off, g.npre, g.src, g.w, c->c_sidx, c->signal, g.src2, g.w2, g.ncnt,
                                               c->counters);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, g.ncnt, n, g.off2));
    k_net_copyback<<<grid_for(n), kBlock, 0, s>>>(n, g.off, g.npre, g.off2, g.src2, g.w2, g.src, g.w, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_offsets<<<grid_for(n + 1), kBlock, 0, s>>>(n, g.off2, g.off, c->counters);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

DelArgs del_args(rl_bh_ctx* c, int type) {
    DelArgs a{};
    a.n = c->n;
    a.net_of = c->net_of;
    a.net_off = c->net_off;
    a.net_seed = c->net_seed;
    a.step = c->d_steps + 1;
    a.deleted = c->neurons.deleted + static_cast<uint64_t>(type) * c->n;
    a.signal = c->signal;
    a.off = c->net.off;
    a.src = c->net.src;
    a.w = c->net.w;
    for (int t = 0; t < 3; t++) {
        a.conn[t] = c->conn[t];
        a.vac[t] = c->vac[t];
    }
    a.counters = c->counters;
    return a;
}

}  // namespace

void net_free(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    dfree(&g.off);
    dfree(&g.src);
    dfree(&g.w);
    dfree(&g.off2);
    dfree(&g.src2);
    dfree(&g.w2);
    dfree(&g.ncnt);
    dfree(&g.npre);
    dfree(&g.ocnt);
    dfree(&g.ooff);
    dfree(&g.olist);
    dfree(&g.out_off);
    dfree(&g.out_cnt);
    dfree(&g.out_tgt);
    dfree(&g.out_w);
    dfree(&g.acc[0]);
    dfree(&g.acc[1]);
    g.cap = 0;
    g.ready = false;
    g.syn_valid = false;
}

int net_init(rl_bh_ctx* c, uint64_t capacity, const uint64_t* in_offsets, const uint32_t* sources,
             const int32_t* weights) {
    const uint64_t n = c->n;
    if (n == 0) {
        return fail(c, RL_ESTATE, "rl_net_init: no neurons (rl_bh_init first)");
    }
    // merge repeated (source, target) pairs and drop zero weights, as NetworkGraph keeps them
    std::vector<uint64_t> off(n + 1, 0);
    std::vector<uint32_t> src;
    std::vector<int32_t> w;
    if (in_offsets != nullptr) {
        if (in_offsets[0] != 0) {
            return fail(c, RL_EINVAL, "rl_net_init: in_offsets[0] must be 0");
        }
        for (uint64_t t = 0; t < n; t++) {
            if (in_offsets[t + 1] < in_offsets[t]) {
                return fail(c, RL_EINVAL, "rl_net_init: in_offsets must be non-decreasing");
            }
        }
        const uint64_t edges = in_offsets[n];
        if (edges > 0 && (sources == nullptr || weights == nullptr)) {
            return fail(c, RL_EINVAL, "rl_net_init: null sources or weights");
        }
        for (uint64_t e = 0; e < edges; e++) {
            if (sources[e] >= n) {
                return fail(c, RL_EINVAL, "rl_net_init: edge source out of range");
            }
        }
        std::vector<std::pair<uint32_t, int64_t>> pairs;
        for (uint64_t t = 0; t < n; t++) {
            pairs.clear();
            for (uint64_t e = in_offsets[t]; e < in_offsets[t + 1]; e++) {
                pairs.emplace_back(sources[e], weights[e]);
            }
            std::sort(pairs.begin(), pairs.end());
            int64_t in_weight = 0;   // the synaptic input sums a neuron's in-weights in 32 bits
            for (size_t i = 0; i < pairs.size();) {
                int64_t sum = 0;
                size_t j = i;
                for (; j < pairs.size() && pairs[j].first == pairs[i].first; j++) {
                    sum += pairs[j].second;
                }
                if (sum != 0) {
                    in_weight += sum < 0 ? -sum : sum;
                    if (sum > INT32_MAX || sum < INT32_MIN || in_weight > INT32_MAX) {
                        return fail(c, RL_EINVAL, "rl_net_init: pair weight or summed in-weight of a neuron out of range");
                    }
                    src.push_back(pairs[i].first);
                    w.push_back(static_cast<int32_t>(sum));
                }
                i = j;
            }
            off[t + 1] = src.size();
        }
    }
    if (src.size() > capacity) {
        return fail(c, RL_EINVAL, "rl_net_init: more edges than the capacity");
    }
    destroy_graph(c);  // a captured period without the graph must be captured again
    net_free(c);
    NetGraph& g = c->net;
    RL_TRY(dalloc(c, &g.off, n + 1));
    RL_TRY(dalloc(c, &g.off2, n + 1));
    RL_TRY(dalloc(c, &g.ncnt, n));
    RL_TRY(dalloc(c, &g.npre, n + 1));
    RL_TRY(dalloc(c, &g.ocnt, n));
    RL_TRY(dalloc(c, &g.ooff, n + 1));
    RL_TRY(dalloc(c, &g.out_off, n + 1));
    RL_TRY(dalloc(c, &g.out_cnt, n));
    RL_TRY(dalloc(c, &g.acc[0], n));
    RL_TRY(dalloc(c, &g.acc[1], n));
    RL_TRY(net_alloc_edges(c, std::max<uint64_t>(capacity, 1)));
    hipStream_t s = c->stream;
    RL_HIP(c, hipMemcpyAsync(g.off, off.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    if (!src.empty()) {
        RL_HIP(c, hipMemcpyAsync(g.src, src.data(), src.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        RL_HIP(c, hipMemcpyAsync(g.w, w.data(), w.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    RL_HIP(c, hipMemsetAsync(g.ocnt, 0, n * sizeof(uint32_t), s));
    RL_HIP(c, hipStreamSynchronize(s));
    g.entries_bound = src.size();
    g.ready = true;
    g.syn_valid = false;
    return RL_OK;
}

int net_get(rl_bh_ctx* c, uint64_t capacity, uint64_t* in_offsets, uint32_t* sources, int32_t* weights,
            uint64_t* num_edges) {
    NetGraph& g = c->net;
    if (!g.ready) {
        return fail(c, RL_ESTATE, "rl_net_get_edges: no network graph (rl_net_init)");
    }
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    std::vector<uint64_t> off(n + 1);
    RL_HIP(c, hipMemcpyAsync(off.data(), g.off, (n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    const uint64_t edges = off[n];
    if (num_edges != nullptr) {
        *num_edges = edges;
    }
    if (in_offsets == nullptr && sources == nullptr && weights == nullptr) {
        return RL_OK;
    }
    if (capacity < edges) {
        return fail(c, RL_EINVAL, "rl_net_get_edges: capacity too small");
    }
    if (in_offsets != nullptr) {
        std::memcpy(in_offsets, off.data(), (n + 1) * sizeof(uint64_t));
    }
    if (edges > 0) {
        if (sources != nullptr) RL_HIP(c, hipMemcpyAsync(sources, g.src, edges * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (weights != nullptr) RL_HIP(c, hipMemcpyAsync(weights, g.w, edges * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    return RL_OK;
}

// Rebuilds the out-edge lists from the graph and the accumulators from the fired status of
// fired_buf[parity]: the state the step kernel expects at that parity.  Allocates nothing and reads
// no count on the host, so a captured period can end with it.
int enqueue_syn_sync(rl_bh_ctx* c, int parity) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    if (!g.ready || n == 0 || !c->neurons.ready) {
        return RL_OK;
    }
    hipStream_t s = c->stream;
    RL_TRY(fill_u32(c, g.out_cnt, 0U, n, nullptr, 0));
    k_out_count<<<kStreamBlocks, kBlock, 0, s>>>(n, g.off, g.src, g.w, g.out_cnt);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, g.out_cnt, n, g.out_off));
    k_out_scatter<<<kStreamBlocks, kBlock, 0, s>>>(n, g.off, g.src, g.w, g.out_off, g.out_cnt, g.out_tgt, g.out_w);
    RL_HIP(c, hipGetLastError());
    RL_TRY(fill_u32(c, reinterpret_cast<uint32_t*>(g.acc[0]), 0U, n, nullptr, 0));
    RL_TRY(fill_u32(c, reinterpret_cast<uint32_t*>(g.acc[1]), 0U, n, nullptr, 0));
    k_syn_push_all<<<kStreamBlocks, kBlock, 0, s>>>(n, c->neurons.fired_buf[parity], g.out_off, g.out_tgt, g.out_w,
                                                    g.acc[parity]);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

int syn_prepare(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    if (!g.ready || g.syn_valid || c->n == 0 || !c->neurons.ready) {
        return RL_OK;
    }
    RL_TRY(enqueue_syn_sync(c, c->neurons.parity));
    g.syn_valid = true;
    return RL_OK;
}

int enqueue_net_append(rl_bh_ctx* c) {
    if (!c->net.ready || c->n == 0) {
        return RL_OK;
    }
    RL_TRY(enqueue_net_count(c));
    return enqueue_net_merge(c);
}

// Grows the edge arrays to hold `need` entries (at least doubling), keeping the current pairs.  A
// captured period refers to the old arrays and is released.
static int net_grow(rl_bh_ctx* c, uint64_t have, uint64_t need) {
    NetGraph& g = c->net;
    hipStream_t s = c->stream;
    destroy_graph(c);
    const uint64_t cap = std::max<uint64_t>(2 * g.cap, need);
    uint32_t* src = nullptr;
    int32_t* w = nullptr;
    RL_TRY(dalloc(c, &src, cap));
    RL_TRY(dalloc(c, &w, cap));
    if (have > 0) {
        RL_HIP(c, hipMemcpyAsync(src, g.src, have * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        RL_HIP(c, hipMemcpyAsync(w, g.w, have * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    dfree(&g.src);
    dfree(&g.w);
    dfree(&g.src2);
    dfree(&g.w2);
    dfree(&g.olist);
    dfree(&g.out_tgt);
    dfree(&g.out_w);
    g.src = src;
    g.w = w;
    g.syn_valid = false;
    RL_TRY(dalloc(c, &g.src2, cap));
    RL_TRY(dalloc(c, &g.w2, cap));
    RL_TRY(dalloc(c, &g.olist, cap));
    RL_TRY(dalloc(c, &g.out_tgt, cap));
    RL_TRY(dalloc(c, &g.out_w, cap));
    g.cap = cap;
    return RL_OK;
}

// Merges the counted synapses (ncnt, npre of the last commit), growing the arrays first when they
// would not fit.
static int net_merge_grow(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    uint64_t have = 0, add = 0;
    RL_HIP(c, hipMemcpyAsync(&have, g.off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(&add, g.npre + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    if (have + add > g.cap) {
        RL_TRY(net_grow(c, have, have + add));
    }
    g.entries_bound = have + add;   // the merge keeps at most the old entries and the new synapses
    return enqueue_net_merge(c);
}

int net_append(rl_bh_ctx* c) {
    if (!c->net.ready || c->n == 0) {
        return RL_OK;
    }
    c->net.syn_valid = false;
    RL_TRY(enqueue_net_count(c));
    return net_merge_grow(c);
}

// A captured append that did not fit left the counts of its synapses (ncnt, npre) and the graph
// untouched: grow and merge them now.
int net_append_retry(rl_bh_ctx* c) {
    if (!c->net.ready || c->n == 0) {
        return RL_OK;
    }
    c->net.syn_valid = false;
    RL_TRY(net_merge_grow(c));
    unsigned long long over = 0;
    RL_HIP(c, hipMemcpyAsync(&over, c->counters + kCtrNetOver, sizeof(over), hipMemcpyDeviceToHost, c->stream));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    if (over != 0) {
        return fail(c, RL_EHIP, "network graph append: still over capacity after growing");
    }
    return RL_OK;
}

// Deletes the synapses of the elements of `type` that the element commit unbound (neurons.deleted).
int enqueue_net_delete(rl_bh_ctx* c, int type) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    if (!g.ready || n == 0) {
        return RL_OK;
    }
    hipStream_t s = c->stream;
    const DelArgs a = del_args(c, type);
    if (type == 0) {
        // out-entries of the sources that delete axons (ocnt is zero between phases)
        k_del_out_count<<<grid_for(n), kBlock, 0, s>>>(n, g.off, g.src, g.w, a.deleted, g.ocnt);
        RL_HIP(c, hipGetLastError());
        RL_TRY(scan_counts(c, g.ocnt, n, g.ooff));
        k_del_out_scatter<<<grid_for(n), kBlock, 0, s>>>(n, g.off, g.src, g.w, a.deleted, g.ooff, g.ocnt, g.olist);
        RL_HIP(c, hipGetLastError());
        k_del_axons<<<grid_for(n), kBlock, 0, s>>>(a, g.ooff, g.olist);
    } else {
        k_del_dendrites<<<grid_for(n), kBlock, 0, s>>>(a, static_cast<uint32_t>(type));
    }
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

}  // namespace rl
