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
__global__ void k_commit_count(uint64_t n, const uint64_t* __restrict__ n_req, const uint32_t* __restrict__ tgt,
                               const uint32_t* __restrict__ src, const uint8_t* __restrict__ sig,
                               const uint32_t* __restrict__ rj, const uint32_t* __restrict__ net_of,
                               const uint64_t* __restrict__ net_off, const uint64_t* __restrict__ net_seed,
                               const uint32_t* __restrict__ step, int counter_order, uint64_t* __restrict__ prio,
                               uint32_t* __restrict__ gcnt, uint32_t* __restrict__ pos, uint8_t* __restrict__ ok) {
    const uint64_t m = *n_req;
    const uint32_t st = *step;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        const uint32_t s = src[i];
        const uint32_t t = tgt[i];
        if (counter_order) {
            const uint32_t g = net_of != nullptr ? net_of[s] : 0U;   // no table: one network
            prio[i] = commit_priority(net_seed[g], s - static_cast<uint32_t>(net_off[g]), st, rj[i]);
        }
        if (s == t || t >= n || s >= n) {   // self request, or an id outside the neurons: fails
            ok[i] = 0;
        } else {
            // the count before this request is its position inside the group
            pos[i] = atomicAdd(&gcnt[2 * static_cast<uint64_t>(t) + (sig[i] & 1U)], 1U);
        }
    }
}

// Step 2, after the prefix sum goff of the group sizes.  A request of a group that has no more
// requests than vacancies is accepted here.  The others go to their group's offset plus the position
// step 1 gave them (an arbitrary order inside the group: step 3 compares priorities, not positions).
// Positions of accepted groups keep the sentinel kUngrouped.
constexpr uint32_t kUngrouped = 0xFFFFFFFFU;

__global__ void k_commit_scatter(uint64_t n, const uint64_t* __restrict__ n_req, const uint32_t* __restrict__ tgt,
                                 const uint32_t* __restrict__ src, const uint8_t* __restrict__ sig,
                                 const uint64_t* __restrict__ prio, const uint64_t* __restrict__ goff,
                                 const uint32_t* __restrict__ vexc, const uint32_t* __restrict__ vinh,
                                 const uint32_t* __restrict__ pos, uint64_t* __restrict__ sprio,
                                 uint32_t* __restrict__ sidx, uint8_t* __restrict__ ok) {
    const uint64_t m = *n_req;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        const uint32_t t = tgt[i];
        const uint32_t sr = src[i];
        if (sr == t || t >= n || sr >= n) {
            continue;
        }
        const uint32_t ty = sig[i] & 1U;
        const uint64_t g = 2 * static_cast<uint64_t>(t) + ty;
        const uint64_t size = goff[g + 1] - goff[g];
        if (size <= static_cast<uint64_t>(ty ? vinh[t] : vexc[t])) {
            ok[i] = 1;
            continue;
        }
        const uint64_t q = goff[g] + pos[i];
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
            if (con != nullptr) {
                con[t] += taken;
            }
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
    // position of each request inside its group: the caller's order is no longer needed once ranked
    uint32_t* gpos = reinterpret_cast<uint32_t*>(c->c_order);
    k_commit_count<<<kStreamBlocks, kBlock, 0, s>>>(n, n_req, c->req_target, c->req_source, c->req_signal, c->req_j,
                                                    c->net_of, c->net_off, c->net_seed, c->d_steps + 1,
                                                    h_order == nullptr ? 1 : 0, c->c_prio, c->c_gcnt, gpos, c->req_ok);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, c->c_gcnt, groups, c->c_goff));
    RL_TRY(fill_u32(c, c->c_sidx, kUngrouped, 0, c->c_goff + groups, c->slots_cap));
    k_commit_scatter<<<kStreamBlocks, kBlock, 0, s>>>(n, n_req, c->req_target, c->req_source, c->req_signal, c->c_prio,
                                                      c->c_goff, c->vac[1], c->vac[2], gpos, c->c_sprio, c->c_sidx,
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

// ---------------------------------------------------------------------------------------------
// Allocation-free commit on caller device buffers (rl_bh_commit_requests_noalloc).
//
// The grouping scratch of the commit (group counts and offsets, scan tile sums, priorities, grouped
// requests, a small header) lives in one file-scope device array, `scratch`, sized for
// kNoallocMaxNeurons neurons and kNoallocMaxRequests requests.  Every GPU that runs the library holds
// its own copy until the process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): the device address of each GPU's copy comes from
// hipGetSymbolAddress with that GPU current, once per GPU, and only that address is used.
#include <mutex>

namespace rl {
namespace {

constexpr uint64_t kNoallocMaxNeurons = 1ULL << 18;
constexpr uint64_t kNoallocMaxRequests = 1ULL << 20;

constexpr size_t align256(size_t b) {
    return (b + 255U) & ~static_cast<size_t>(255U);
}

// byte offsets inside the array
constexpr size_t kNaHdr = 0;   // NoallocHeader
constexpr size_t kNaGcnt = 256;
constexpr size_t kNaGoff = kNaGcnt + align256(2 * kNoallocMaxNeurons * sizeof(uint32_t));
constexpr size_t kNaPart = kNaGoff + align256((2 * kNoallocMaxNeurons + 1) * sizeof(uint64_t));
constexpr size_t kNaPrio = kNaPart + align256(((2 * kNoallocMaxNeurons + 4095) / 4096 + 1) * sizeof(uint64_t));
constexpr size_t kNaSprio = kNaPrio + align256(kNoallocMaxRequests * sizeof(uint64_t));
constexpr size_t kNaSidx = kNaSprio + align256(kNoallocMaxRequests * sizeof(uint64_t));
constexpr size_t kNaPos = kNaSidx + align256(kNoallocMaxRequests * sizeof(uint32_t));
constexpr size_t kNaBytes = kNaPos + align256(kNoallocMaxRequests * sizeof(uint32_t));

struct NoallocHeader {
    uint64_t m;                    // requests of the call (the kernels read the count from here)
    uint64_t net_off;              // 0: the requests' source ids are the ids in the network
    uint64_t seed;                 // network seed of the counter priorities
    unsigned long long created;    // accepted requests
    uint32_t step;                 // plasticity update addressing the priorities
};
static_assert(sizeof(NoallocHeader) <= 256, "header slot");

__device__ __align__(256) unsigned char scratch[kNaBytes];

__global__ void k_noalloc_setup(NoallocHeader* h, uint64_t m, uint64_t seed, uint32_t step) {
    if (threadIdx.x == 0) {
        h->m = m;
        h->net_off = 0;
        h->seed = seed;
        h->created = 0;
        h->step = step;
    }
}

// Step 4, axon side, for requests in any order: every accepted request takes one vacant axon of its
// source (the caller's requests of one source never outnumber its vacant axons).
__global__ void k_commit_axons_by_request(const uint64_t* __restrict__ n_req, const uint32_t* __restrict__ src,
                                          const uint8_t* __restrict__ ok, uint32_t* __restrict__ vax,
                                          uint32_t* __restrict__ cax) {
    const uint64_t m = *n_req;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < m; i += stride) {
        if (ok[i] != 0) {
            const uint32_t s = src[i];
            atomicSub(&vax[s], 1U);
            if (cax != nullptr) {
                atomicAdd(&cax[s], 1U);
            }
        }
    }
}

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mutex;
unsigned char* g_noalloc_ptr[kNoallocMaxDevices] = {};

// Device address of the current GPU's copy of `scratch`, looked up on the first call on that GPU.
int noalloc_scratch(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_global_error("rl_bh_commit_requests_noalloc: hipGetDevice failed");
        return RL_EHIP;
    }
    if (dev < 0 || dev >= kNoallocMaxDevices) {
        set_global_error("rl_bh_commit_requests_noalloc: device ordinal beyond 63");
        return RL_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_global_error("rl_bh_commit_requests_noalloc: hipGetSymbolAddress of the scratch array failed");
            return RL_EHIP;
        }
        g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return RL_OK;
}

#define RL_NA_HIP(expr)                                                                   \
    do {                                                                                  \
        const hipError_t rl_e_ = (expr);                                                  \
        if (rl_e_ != hipSuccess) {                                                        \
            set_global_error("rl_bh_commit_requests_noalloc: HIP launch or copy failed"); \
            return RL_EHIP;                                                               \
        }                                                                                 \
    } while (0)

}  // namespace
}  // namespace rl

extern "C" int rl_bh_commit_requests_noalloc(uint64_t n, uint64_t m, const uint32_t* d_target,
                                             const uint32_t* d_source, const uint8_t* d_signal,
                                             const uint32_t* d_axon_ordinal, uint64_t network_seed, uint32_t step,
                                             uint32_t* d_vacant_axons, uint32_t* d_vacant_excitatory_dendrites,
                                             uint32_t* d_vacant_inhibitory_dendrites, uint32_t* d_connected_axons,
                                             uint32_t* d_connected_excitatory_dendrites,
                                             uint32_t* d_connected_inhibitory_dendrites, uint8_t* d_ok,
                                             void* hip_stream, uint64_t* created) {
    using namespace rl;
    if (created != nullptr) {
        *created = 0;
    }
    if (n > kNoallocMaxNeurons || m > kNoallocMaxRequests) {
        set_global_error("rl_bh_commit_requests_noalloc: more than 262144 neurons or 1048576 requests "
                         "(the size of the static scratch array)");
        return RL_ELIMIT;
    }
    if (m == 0) {
        return RL_OK;
    }
    if (d_target == nullptr || d_source == nullptr || d_signal == nullptr || d_axon_ordinal == nullptr ||
        d_vacant_axons == nullptr || d_vacant_excitatory_dendrites == nullptr ||
        d_vacant_inhibitory_dendrites == nullptr || d_ok == nullptr || n == 0) {
        set_global_error("rl_bh_commit_requests_noalloc: null request, vacancy or result array, or n = 0");
        return RL_EINVAL;
    }
    unsigned char* base = nullptr;
    RL_TRY(noalloc_scratch(&base));
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    auto* hdr = reinterpret_cast<NoallocHeader*>(base + kNaHdr);
    auto* gcnt = reinterpret_cast<uint32_t*>(base + kNaGcnt);
    auto* goff = reinterpret_cast<uint64_t*>(base + kNaGoff);
    auto* part = reinterpret_cast<uint64_t*>(base + kNaPart);
    auto* prio = reinterpret_cast<uint64_t*>(base + kNaPrio);
    auto* sprio = reinterpret_cast<uint64_t*>(base + kNaSprio);
    auto* sidx = reinterpret_cast<uint32_t*>(base + kNaSidx);
    auto* gpos = reinterpret_cast<uint32_t*>(base + kNaPos);
    const uint64_t groups = 2 * n;
    // the array keeps the last call's contents: everything read before it is written is set here
    k_noalloc_setup<<<1, 64, 0, s>>>(hdr, m, network_seed, step);
    RL_NA_HIP(hipGetLastError());
    RL_NA_HIP(fill_u32_on(s, gcnt, 0U, groups, nullptr, 0));
    k_commit_count<<<kStreamBlocks, kBlock, 0, s>>>(n, &hdr->m, d_target, d_source, d_signal, d_axon_ordinal, nullptr,
                                                    &hdr->net_off, &hdr->seed, &hdr->step, 1, prio, gcnt, gpos, d_ok);
    RL_NA_HIP(hipGetLastError());
    RL_NA_HIP(scan_counts_on(s, gcnt, groups, goff, part));
    RL_NA_HIP(fill_u32_on(s, sidx, kUngrouped, 0, goff + groups, kNoallocMaxRequests));
    k_commit_scatter<<<kStreamBlocks, kBlock, 0, s>>>(n, &hdr->m, d_target, d_source, d_signal, prio, goff,
                                                      d_vacant_excitatory_dendrites, d_vacant_inhibitory_dendrites,
                                                      gpos, sprio, sidx, d_ok);
    RL_NA_HIP(hipGetLastError());
    k_commit_rank<<<kStreamBlocks, kBlock, 0, s>>>(goff + groups, goff, sprio, sidx, d_target, d_signal,
                                                   d_vacant_excitatory_dendrites, d_vacant_inhibitory_dendrites, d_ok);
    RL_NA_HIP(hipGetLastError());
    k_commit_groups<<<kStreamBlocks, kBlock, 0, s>>>(groups, goff, d_vacant_excitatory_dendrites,
                                                     d_vacant_inhibitory_dendrites, d_connected_excitatory_dendrites,
                                                     d_connected_inhibitory_dendrites, &hdr->created);
    RL_NA_HIP(hipGetLastError());
    k_commit_axons_by_request<<<kStreamBlocks, kBlock, 0, s>>>(&hdr->m, d_source, d_ok, d_vacant_axons,
                                                               d_connected_axons);
    RL_NA_HIP(hipGetLastError());
    if (created != nullptr) {
        unsigned long long hc = 0;
        RL_NA_HIP(hipMemcpyAsync(&hc, &hdr->created, sizeof(hc), hipMemcpyDeviceToHost, s));
        RL_NA_HIP(hipStreamSynchronize(s));
        *created = hc;
    }
    return RL_OK;
}
