// Request conflict resolution and vacancy commit on the device.
//
// The reference shuffles the request indices and then walks them serially: a request fails when it
// targets its own source or when the target has no vacant dendrite of the requested type left,
// otherwise it takes one (ForwardConnector::process_requests, source/algorithm/Connector.cpp:33-121);
// every success then takes one vacant axon of its source (process_responses, :123-197).  Requests for
// different (target, dendrite type) pairs never interact, so the serial walk equals: sort the
// requests by (target, type, position in the processing order) and let one lane per (target, type)
// group walk its segment.  The processing order is the caller's permutation (parity with a given
// shuffle) or a counter-based 64-bit priority per request (ties by request index).
#include <hipcub/hipcub.hpp>

#include "rl_device.h"
#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

// Counter-based commit priority of request i, addressed like the draws: the network's seed and the
// source's id inside its network.
__global__ void k_priority(uint64_t m, const uint32_t* __restrict__ src, const uint32_t* __restrict__ rj,
                           const uint32_t* __restrict__ net_of, const uint64_t* __restrict__ net_off,
                           const uint64_t* __restrict__ net_seed, uint32_t step, uint64_t* __restrict__ prio,
                           uint32_t* __restrict__ idx) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < m) {
        const uint32_t s = src[i];
        const uint32_t g = net_of[s];
        prio[i] = commit_priority(net_seed[g], s - static_cast<uint32_t>(net_off[g]), step, rj[i]);
        idx[i] = static_cast<uint32_t>(i);
    }
}

__global__ void k_rank_from_order(uint64_t m, const uint64_t* __restrict__ order, uint64_t* __restrict__ prio,
                                  uint32_t* __restrict__ idx, int* __restrict__ err) {
    const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (k < m) {
        const uint64_t i = order[k];
        if (i >= m) {
            atomicOr(err, 1);
            return;
        }
        prio[i] = k;
        idx[k] = static_cast<uint32_t>(k);
    }
}

__global__ void k_group_keys(uint64_t m, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ tgt,
                             const uint8_t* __restrict__ sig, uint32_t* __restrict__ g) {
    const uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (q < m) {
        const uint32_t i = idx[q];
        g[q] = (tgt[i] << 1U) | (sig[i] & 1U);
    }
}

// One lane per (target, type) group walks its requests in processing order.
__global__ void k_resolve(uint64_t m, const uint32_t* __restrict__ g, const uint32_t* __restrict__ idx,
                          const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ src, uint32_t* __restrict__ vexc,
                          uint32_t* __restrict__ vinh, uint32_t* __restrict__ cexc, uint32_t* __restrict__ cinh,
                          uint8_t* __restrict__ ok, unsigned long long* __restrict__ created) {
    const uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (q >= m) {
        return;
    }
    const uint32_t key = g[q];
    if (q > 0 && g[q - 1] == key) {
        return;
    }
    const uint32_t target = key >> 1U;
    uint32_t* vac = (key & 1U) ? vinh : vexc;
    uint32_t* con = (key & 1U) ? cinh : cexc;
    uint32_t left = vac[target];
    uint32_t taken = 0;
    for (uint64_t r = q; r < m && g[r] == key; r++) {
        const uint32_t i = idx[r];
        uint8_t success = 0;
        if (src[i] != target && left > 0) {
            left--;
            taken++;
            success = 1;
        }
        ok[i] = success;
    }
    vac[target] = left;
    con[target] += taken;
    if (taken > 0) {
        atomicAdd(created, static_cast<unsigned long long>(taken));
    }
}

__global__ void k_axons(uint64_t m, const uint8_t* __restrict__ ok, const uint32_t* __restrict__ src,
                        uint32_t* __restrict__ vaxon, uint32_t* __restrict__ caxon) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < m && ok[i]) {
        atomicSub(&vaxon[src[i]], 1U);
        atomicAdd(&caxon[src[i]], 1U);
    }
}

}  // namespace

int commit(rl_bh_ctx* c, uint32_t step, const uint64_t* h_order, uint64_t* created) {
    if (!c->searched_valid) {
        return fail(c, RL_ESTATE, "commit called before find_target_neurons");
    }
    const uint64_t m = c->nreq;
    hipStream_t s = c->stream;
    if (created != nullptr) {
        *created = 0;
    }
    if (m == 0) {
        return RL_OK;
    }
    if (m > c->commit_cap) {
        RL_TRY(dalloc(c, &c->c_prio, m));
        RL_TRY(dalloc(c, &c->c_prio2, m));
        RL_TRY(dalloc(c, &c->c_order, m));
        RL_TRY(dalloc(c, &c->c_idx, m));
        RL_TRY(dalloc(c, &c->c_idx2, m));
        RL_TRY(dalloc(c, &c->c_g, m));
        RL_TRY(dalloc(c, &c->c_g2, m));
        c->commit_cap = m;
    }
    uint64_t* prio = c->c_prio;
    uint64_t* prio2 = c->c_prio2;
    uint64_t* order = c->c_order;
    uint32_t* idx = c->c_idx;
    uint32_t* idx2 = c->c_idx2;
    uint32_t* g = c->c_g;
    uint32_t* g2 = c->c_g2;
    int* err = c->flags;
    unsigned long long* cnt = c->counters + 4;
    int status = RL_OK;
    do {
#define RL_STEP(expr)                            \
    {                                            \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) {                  \
            status = hip_fail(c, e_, #expr);     \
            break;                               \
        }                                        \
    }
        RL_STEP(hipMemsetAsync(err, 0, sizeof(int), s));
        RL_STEP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
        if (h_order != nullptr) {
            RL_STEP(hipMemcpyAsync(order, h_order, m * sizeof(uint64_t), hipMemcpyHostToDevice, s));
            RL_STEP(hipMemsetAsync(prio, 0xFF, m * sizeof(uint64_t), s));
            k_rank_from_order<<<grid_for(m), kBlock, 0, s>>>(m, order, prio, idx, err);
        } else {
            k_priority<<<grid_for(m), kBlock, 0, s>>>(m, c->req_source, c->req_j, c->net_of, c->net_off, c->net_seed,
                                                      step, prio, idx);
        }
        RL_STEP(hipGetLastError());
        size_t tb = 0, tb2 = 0;
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, prio, prio2, idx, idx2, static_cast<int>(m), 0, 64, s));
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, g, g2, idx2, idx, static_cast<int>(m), 0, 32, s));
        if ((status = ensure_tmp(c, std::max(tb, tb2))) != RL_OK) {
            break;
        }
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tb, prio, prio2, idx, idx2, static_cast<int>(m), 0, 64, s));
        k_group_keys<<<grid_for(m), kBlock, 0, s>>>(m, idx2, c->req_target, c->req_signal, g);
        RL_STEP(hipGetLastError());
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tb2, g, g2, idx2, idx, static_cast<int>(m), 0, 32, s));
        k_resolve<<<grid_for(m), kBlock, 0, s>>>(m, g2, idx, c->req_target, c->req_source, c->vac[1], c->vac[2],
                                                  c->conn[1], c->conn[2], c->req_ok, cnt);
        RL_STEP(hipGetLastError());
        k_axons<<<grid_for(m), kBlock, 0, s>>>(m, c->req_ok, c->req_source, c->vac[0], c->conn[0]);
        RL_STEP(hipGetLastError());
        int herr = 0;
        unsigned long long hc = 0;
        RL_STEP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
        RL_STEP(hipMemcpyAsync(&hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
        RL_STEP(hipStreamSynchronize(s));
        if (herr) {
            status = fail(c, RL_EINVAL, "order is not a permutation of the request indices");
            break;
        }
        if (created != nullptr) {
            *created = hc;
        }
#undef RL_STEP
    } while (false);
    return status;
}

}  // namespace rl
