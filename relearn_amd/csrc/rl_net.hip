// Network graph of the synaptic input on the device.
//
// The synaptic input of a neuron (SynapticEquallyWeightedActivityInputCPU::update_local_input,
// source/neurons/input/SynapticEquallyWeightedActivityInputCPU.cpp:20-50) sums the weights of its
// in-edges (NetworkGraph::get_local_in_edges, static and plastic) whose source fired in the last
// step.  Only the multiset of (source, weight) per target matters, so the device keeps a CSR by
// target with one entry per edge; the reference merges repeated (source, target) pairs into one edge
// whose weight is the sum, which the input sums to the same value.  Every commit appends the synapses
// it created (Neurons::create_synapses adds them to the graph, source/neurons/Neurons.cpp:440-477),
// weight +1 for an excitatory source and -1 for an inhibitory one
// (ForwardConnector::process_requests, source/algorithm/Connector.cpp:33-121).
//
// The append is a counting sort merge: count the new in-edges per target, prefix-sum them, move
// every target's old edges to their shifted place in the second buffer, scatter the new ones after
// them, copy the merged graph back.  Counts come from device memory and nothing allocates, so the
// append can be captured with the commit.  An append that would exceed the capacity appends nothing
// and raises a device flag.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

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

__global__ void k_net_check(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ npre,
                            uint64_t cap, unsigned long long* __restrict__ counters) {
    if (threadIdx.x == 0) {
        const uint64_t add = npre[n];
        const bool over = off[n] + add > cap;
        counters[kCtrNetOver] = over ? 1ULL : 0ULL;
        counters[kCtrNetNew] = over ? 0ULL : add;
        if (over) {
            counters[kCtrSticky] = 1ULL;
        }
    }
}

// Target t's old edges move from off[t] to off[t] + npre[t] in the second buffer.
__global__ void k_net_move(uint64_t n, const uint64_t* __restrict__ off, const uint64_t* __restrict__ npre,
                           const uint32_t* __restrict__ src, const int32_t* __restrict__ w, uint64_t* __restrict__ off2,
                           uint32_t* __restrict__ src2, int32_t* __restrict__ w2,
                           const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t <= n; t += stride) {
        const uint64_t o = off[t];
        const uint64_t o2 = o + npre[t];
        off2[t] = o2;
        if (t < n) {
            const uint64_t d = off[t + 1] - o;
            for (uint64_t k = 0; k < d; k++) {
                src2[o2 + k] = src[o + k];
                w2[o2 + k] = w[o + k];
            }
        }
    }
}

// New edges go after the target's old ones; their order inside a target is arbitrary (the input
// sums integer weights, which is exact in any order).  ncnt returns to zero.
__global__ void k_net_scatter(const uint64_t* __restrict__ n_req, const uint8_t* __restrict__ ok,
                              const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ rsrc,
                              const uint8_t* __restrict__ rsig, const uint64_t* __restrict__ off,
                              const uint64_t* __restrict__ off2, uint32_t* __restrict__ ncnt, uint32_t* __restrict__ src2,
                              int32_t* __restrict__ w2, const unsigned long long* __restrict__ counters) {
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
        const uint64_t pos = off2[t] + (off[t + 1] - off[t]) + atomicSub(&ncnt[t], 1U) - 1U;
        src2[pos] = rsrc[i];
        w2[pos] = rsig[i] != 0 ? -1 : 1;
    }
}

__global__ void k_net_copyback(uint64_t n, const uint64_t* __restrict__ off2, const uint32_t* __restrict__ src2,
                               const int32_t* __restrict__ w2, uint64_t* __restrict__ off, uint32_t* __restrict__ src,
                               int32_t* __restrict__ w, const unsigned long long* __restrict__ counters) {
    if (counters[kCtrNetOver] != 0) {
        return;
    }
    const uint64_t edges = off2[n];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t t0 = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    for (uint64_t t = t0; t <= n; t += stride) {
        off[t] = off2[t];
    }
    for (uint64_t e = t0; e < edges; e += stride) {
        src[e] = src2[e];
        w[e] = w2[e];
    }
}

int net_alloc_edges(rl_bh_ctx* c, uint64_t cap) {
    NetGraph& g = c->net;
    RL_TRY(dalloc(c, &g.src, cap));
    RL_TRY(dalloc(c, &g.w, cap));
    RL_TRY(dalloc(c, &g.src2, cap));
    RL_TRY(dalloc(c, &g.w2, cap));
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

int enqueue_net_merge(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    k_net_check<<<1, 64, 0, s>>>(n, g.off, g.npre, g.cap, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_move<<<grid_for(n + 1), kBlock, 0, s>>>(n, g.off, g.npre, g.src, g.w, g.off2, g.src2, g.w2, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_scatter<<<kStreamBlocks, kBlock, 0, s>>>(c->req_off + n, c->req_ok, c->req_target, c->req_source,
                                                   c->req_signal, g.off, g.off2, g.ncnt, g.src2, g.w2, c->counters);
    RL_HIP(c, hipGetLastError());
    k_net_copyback<<<kStreamBlocks, kBlock, 0, s>>>(n, g.off2, g.src2, g.w2, g.off, g.src, g.w, c->counters);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
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
    g.cap = 0;
    g.ready = false;
}

int net_init(rl_bh_ctx* c, uint64_t capacity, const uint64_t* in_offsets, const uint32_t* sources,
             const int32_t* weights) {
    const uint64_t n = c->n;
    if (n == 0) {
        return fail(c, RL_ESTATE, "rl_net_init: no neurons (rl_bh_init first)");
    }
    uint64_t edges = 0;
    if (in_offsets != nullptr) {
        if (in_offsets[0] != 0) {
            return fail(c, RL_EINVAL, "rl_net_init: in_offsets[0] must be 0");
        }
        for (uint64_t t = 0; t < n; t++) {
            if (in_offsets[t + 1] < in_offsets[t]) {
                return fail(c, RL_EINVAL, "rl_net_init: in_offsets must be non-decreasing");
            }
        }
        edges = in_offsets[n];
        if (edges > 0 && (sources == nullptr || weights == nullptr)) {
            return fail(c, RL_EINVAL, "rl_net_init: null sources or weights");
        }
        for (uint64_t e = 0; e < edges; e++) {
            if (sources[e] >= n) {
                return fail(c, RL_EINVAL, "rl_net_init: edge source out of range");
            }
        }
    }
    if (edges > capacity) {
        return fail(c, RL_EINVAL, "rl_net_init: more edges than the capacity");
    }
    destroy_graph(c);  // a captured period without the append must be captured again
    net_free(c);
    NetGraph& g = c->net;
    RL_TRY(dalloc(c, &g.off, n + 1));
    RL_TRY(dalloc(c, &g.off2, n + 1));
    RL_TRY(dalloc(c, &g.ncnt, n));
    RL_TRY(dalloc(c, &g.npre, n + 1));
    RL_TRY(net_alloc_edges(c, std::max<uint64_t>(capacity, 1)));
    hipStream_t s = c->stream;
    if (in_offsets != nullptr) {
        RL_HIP(c, hipMemcpyAsync(g.off, in_offsets, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        if (edges > 0) {
            RL_HIP(c, hipMemcpyAsync(g.src, sources, edges * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            RL_HIP(c, hipMemcpyAsync(g.w, weights, edges * sizeof(int32_t), hipMemcpyHostToDevice, s));
        }
    } else {
        RL_HIP(c, hipMemsetAsync(g.off, 0, (n + 1) * sizeof(uint64_t), s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    g.ready = true;
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

int enqueue_net_append(rl_bh_ctx* c) {
    if (!c->net.ready || c->n == 0) {
        return RL_OK;
    }
    RL_TRY(enqueue_net_count(c));
    return enqueue_net_merge(c);
}

int net_append(rl_bh_ctx* c) {
    NetGraph& g = c->net;
    if (!g.ready || c->n == 0) {
        return RL_OK;
    }
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    RL_TRY(enqueue_net_count(c));
    uint64_t have = 0, add = 0;
    RL_HIP(c, hipMemcpyAsync(&have, g.off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(&add, g.npre + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    if (have + add > g.cap) {
        // grow: keep the current edges, double the capacity; a captured period refers to the old
        // arrays and is released
        destroy_graph(c);
        const uint64_t cap = std::max<uint64_t>(2 * g.cap, have + add);
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
        g.src = src;
        g.w = w;
        RL_TRY(dalloc(c, &g.src2, cap));
        RL_TRY(dalloc(c, &g.w2, cap));
        g.cap = cap;
    }
    return enqueue_net_merge(c);
}

}  // namespace rl
