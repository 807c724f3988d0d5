// Request conflict resolution and vacancy commit on the device.
//
// The reference shuffles the request indices and then walks them serially: a request fails when it
// targets its own source or when the target has no vacant dendrite of the requested type left,
// otherwise it takes one (ForwardConnector::process_requests, source/algorithm/Connector.cpp:33-121);
// every success then takes one vacant axon of its source (process_responses, :123-197).  Requests of
// different (target, dendrite type) groups never interact and a self request takes nothing, so the
// serial walk accepts, in every group with v vacant dendrites, exactly the first v of its non-self
// requests in processing order.  On the device:
//   1. self requests fail; the others are counted per group (and get their priority);
//   2. a group with no more requests than vacancies accepts all of them (about half of the requests
//      at 1M-16M neurons); the requests of the other groups are grouped by a counting sort (prefix
//      sum of the group sizes, scatter);
//   3. each grouped request counts the members of its group that come before it in processing
//      order, stopping at v, and is accepted when fewer than v do;
//   4. each group takes min(v, size) vacancies, each source one vacant axon per accepted request.
// The processing order is the caller's permutation (parity with a given shuffle) or a counter-based
// 64-bit priority per request, ties by request index.  Nothing here sorts by priority, allocates or
// synchronises with the host; the request count is read from device memory.
#include "rl_device.h"
#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

__global__ void k_rank_from_order(uint64_t m, const uint64_t* __restrict__ order, uint64_t* __restrict__ prio) {
    const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (k < m) {
        const uint64_t i = order[k];
        if (i < m) {
            prio[i] = k;
        }
    }
}

// Step 1.  The counter-based priority of request i is addressed like the draws: the network's seed,
// the source's id inside its network, the update index and the axon ordinal.
__global__ void k_commit_count(const uint64_t* __restrict__ n_req, const uint32_t* __restrict__ tgt,
                               const uint32_t* __restrict__ src, const uint8_t* __restrict__ sig,
                               const uint32_t* __restrict__ rj, const uint32_t* __restrict__ net_of,
                               const uint64_t* __restrict__ net_off, const uint64_t* __restrict__ net_seed,
                               const uint32_t* __restrict__ step, int counter_order, uint64_t* __restrict__ prio,
                               uint32_t* __restrict__ gcnt, uint8_t* __restrict__ ok) {
    const uint64_t m = *n_req;
    const uint32_t st = *step;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        const uint32_t s = src[i];
        const uint32_t t = tgt[i];
        if (counter_order) {
            const uint32_t g = net_of[s];
            prio[i] = commit_priority(net_seed[g], s - static_cast<uint32_t>(net_off[g]), st, rj[i]);
        }
        if (s == t) {
            ok[i] = 0;
        } else {
            atomicAdd(&gcnt[2 * static_cast<uint64_t>(t) + (sig[i] & 1U)], 1U);
        }
    }
}

// Step 2, after the prefix sum goff of the group sizes.  A request of a group that has no more
// requests than vacancies is accepted here.  The others fill their group from its end; the order
// inside a group is arbitrary (step 3 compares priorities, not positions).  Positions of accepted
// groups keep the sentinel kUngrouped.
constexpr uint32_t kUngrouped = 0xFFFFFFFFU;

__global__ void k_commit_scatter(const uint64_t* __restrict__ n_req, const uint32_t* __restrict__ tgt,
                                 const uint32_t* __restrict__ src, const uint8_t* __restrict__ sig,
                                 const uint64_t* __restrict__ prio, const uint64_t* __restrict__ goff,
                                 const uint32_t* __restrict__ vexc, const uint32_t* __restrict__ vinh,
                                 uint32_t* __restrict__ gcnt, uint64_t* __restrict__ sprio, uint32_t* __restrict__ sidx,
                                 uint8_t* __restrict__ ok) {
    const uint64_t m = *n_req;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        const uint32_t t = tgt[i];
        if (src[i] == t) {
            continue;
        }
        const uint32_t ty = sig[i] & 1U;
        const uint64_t g = 2 * static_cast<uint64_t>(t) + ty;
        const uint64_t size = goff[g + 1] - goff[g];
        if (size <= static_cast<uint64_t>(ty ? vinh[t] : vexc[t])) {
            ok[i] = 1;
            continue;
        }
        const uint64_t q = goff[g] + atomicSub(&gcnt[g], 1U) - 1U;
        sprio[q] = prio[i];
        sidx[q] = static_cast<uint32_t>(i);
    }
}

// Step 3: request i (at grouped position q) is accepted iff fewer than v members of its group have a
// smaller (priority, index).  The count stops at v, so a group of size L costs O(L v log L) in all.
__global__ void k_commit_rank(const uint64_t* __restrict__ n_grouped, const uint64_t* __restrict__ goff,
                              const uint64_t* __restrict__ sprio, const uint32_t* __restrict__ sidx,
                              const uint32_t* __restrict__ tgt, const uint8_t* __restrict__ sig,
                              const uint32_t* __restrict__ vexc, const uint32_t* __restrict__ vinh,
                              uint8_t* __restrict__ ok) {
    const uint64_t mg = *n_grouped;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t q = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; q < mg; q += stride) {
        const uint32_t i = sidx[q];
        if (i == kUngrouped) {
            continue;   // a group accepted whole in step 2
        }
        const uint64_t p = sprio[q];
        const uint32_t t = tgt[i];
        const uint32_t ty = sig[i] & 1U;
        const uint64_t g = 2 * static_cast<uint64_t>(t) + ty;
        const uint64_t lo = goff[g], hi = goff[g + 1];
        const uint32_t v = ty ? vinh[t] : vexc[t];
        uint32_t before = 0;
        for (uint64_t r = lo; r < hi && before < v; r++) {
            const uint64_t pr = sprio[r];
            before += (pr < p || (pr == p && sidx[r] < i)) ? 1U : 0U;
        }
        ok[i] = before < v ? 1 : 0;
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, kWave);
    }
    return v;
}

// Step 4, dendrite side: group g = 2 target + type takes min(v, size) vacancies.
__global__ void k_commit_groups(uint64_t groups, const uint64_t* __restrict__ goff, uint32_t* __restrict__ vexc,
                                uint32_t* __restrict__ vinh, uint32_t* __restrict__ cexc, uint32_t* __restrict__ cinh,
                                unsigned long long* __restrict__ created) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    unsigned long long mine = 0;
    for (uint64_t g = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; g < groups; g += stride) {
        const uint64_t size = goff[g + 1] - goff[g];
        if (size == 0) {
            continue;
        }
        const uint64_t t = g >> 1U;
        uint32_t* vac = (g & 1U) ? vinh : vexc;
        uint32_t* con = (g & 1U) ? cinh : cexc;
        const uint32_t v = vac[t];
        const uint32_t taken = size < v ? static_cast<uint32_t>(size) : v;
        if (taken > 0) {
            vac[t] = v - taken;
            con[t] += taken;
            mine += taken;
        }
    }
    mine = wave_sum(mine);
    if ((threadIdx.x & (kWave - 1)) == 0 && mine != 0) {
        atomicAdd(created, mine);
    }
}

// Step 4, axon side: a source's requests are contiguous (found[i] of them at req_off[i]).
__global__ void k_commit_axons(uint64_t n, const uint32_t* __restrict__ found, const uint64_t* __restrict__ req_off,
                               const uint8_t* __restrict__ ok, uint32_t* __restrict__ vax, uint32_t* __restrict__ cax) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t f = found[i];
    if (f == 0) {
        return;
    }
    const uint64_t o = req_off[i];
    uint32_t acc = 0;
    for (uint32_t j = 0; j < f; j++) {
        acc += ok[o + j];
    }
    if (acc != 0) {
        vax[i] -= acc;
        cax[i] += acc;
    }
}

}  // namespace

// The commit of the requests of the last search (count in req_off[n]) on the context's stream.  With
// h_order (host, c->nreq entries, a permutation checked by the caller) the requests are processed in
// that order; otherwise by counter-based priority with the update index d_steps[1].  h_order is the
// only host input and is not allowed while capturing.
int enqueue_commit(rl_bh_ctx* c, const uint64_t* h_order) {
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    if (n == 0) {
        return RL_OK;
    }
    if (h_order != nullptr && c->capturing) {
        return fail(c, RL_ESTATE, "an explicit commit order cannot be captured");
    }
    const uint64_t groups = 2 * n;
    RL_TRY(fill_u32(c, c->c_gcnt, 0U, groups, nullptr, 0));
    RL_TRY(fill_u32(c, reinterpret_cast<uint32_t*>(c->counters + kCtrCreated), 0U, 2, nullptr, 0));
    if (h_order != nullptr && c->nreq > 0) {
        const uint64_t m = c->nreq;
        RL_HIP(c, hipMemcpyAsync(c->c_order, h_order, m * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        k_rank_from_order<<<grid_for(m), kBlock, 0, s>>>(m, c->c_order, c->c_prio);
        RL_HIP(c, hipGetLastError());
    }
    const uint64_t* n_req = c->req_off + n;
    k_commit_count<<<kStreamBlocks, kBlock, 0, s>>>(n_req, c->req_target, c->req_source, c->req_signal, c->req_j,
                                                    c->net_of, c->net_off, c->net_seed, c->d_steps + 1,
                                                    h_order == nullptr ? 1 : 0, c->c_prio, c->c_gcnt, c->req_ok);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, c->c_gcnt, groups, c->c_goff));
    RL_TRY(fill_u32(c, c->c_sidx, kUngrouped, 0, c->c_goff + groups, c->slots_cap));
    k_commit_scatter<<<kStreamBlocks, kBlock, 0, s>>>(n_req, c->req_target, c->req_source, c->req_signal, c->c_prio,
                                                      c->c_goff, c->vac[1], c->vac[2], c->c_gcnt, c->c_sprio, c->c_sidx,
                                                      c->req_ok);
    RL_HIP(c, hipGetLastError());
    k_commit_rank<<<kStreamBlocks, kBlock, 0, s>>>(c->c_goff + groups, c->c_goff, c->c_sprio, c->c_sidx,
                                                   c->req_target, c->req_signal, c->vac[1], c->vac[2], c->req_ok);
    RL_HIP(c, hipGetLastError());
    k_commit_groups<<<kStreamBlocks, kBlock, 0, s>>>(groups, c->c_goff, c->vac[1], c->vac[2], c->conn[1], c->conn[2],
                                                     c->counters + kCtrCreated);
    RL_HIP(c, hipGetLastError());
    k_commit_axons<<<grid_for(n), kBlock, 0, s>>>(n, c->found, c->req_off, c->req_ok, c->vac[0], c->conn[0]);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

int commit(rl_bh_ctx* c, uint32_t step, const uint64_t* h_order, uint64_t* created) {
    if (!c->searched_valid) {
        return fail(c, RL_ESTATE, "commit called before find_target_neurons");
    }
    if (c->committed) {
        return fail(c, RL_ESTATE, "commit: the requests of the last search were already committed");
    }
    if (created != nullptr) {
        *created = 0;
    }
    if (c->nreq == 0) {
        return RL_OK;
    }
    RL_TRY(set_steps(c, 0, step));
    RL_TRY(enqueue_commit(c, h_order));
    c->committed = true;
    unsigned long long hc = 0;
    RL_HIP(c, hipMemcpyAsync(&hc, c->counters + kCtrCreated, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    if (created != nullptr) {
        *created = hc;
    }
    return RL_OK;
}

}  // namespace rl
