// Exclusive prefix sums and fills for the update pipeline.  Sizes come from the host (fixed per
// context) or are read from device memory by the kernels, and nothing here issues a memset, an
// allocation or a host synchronisation, so the pipeline that uses them can be captured into a HIP
// graph and replayed.  The sums are deterministic: the tile partition and the reduction order do not
// depend on timing.
#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;                      // counts per thread
constexpr uint64_t kScanTile = kScanBlock * kScanItems;
constexpr int kTopBlock = 1024;                     // the single block that scans the tile sums
constexpr int kTopItems = 8;
constexpr int kWave = 64;

// Thread t of tile b owns counts [b * 4096 + 16 t, + 16); counts at or beyond n read as 0.
__device__ __forceinline__ void load_counts(const uint32_t* __restrict__ in, uint64_t first, uint64_t n,
                                            uint32_t (&v)[kScanItems]) {
    if (first + kScanItems <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(in + first);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; k++) {
            const uint4 w = p[k];
            v[4 * k] = w.x;
            v[4 * k + 1] = w.y;
            v[4 * k + 2] = w.z;
            v[4 * k + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            v[k] = first + k < n ? in[first + k] : 0U;
        }
    }
}

// Exclusive scan of one value per thread over the block; *total receives the block's sum.
template <int kThreads>
__device__ __forceinline__ uint64_t block_exclusive(uint64_t x, uint64_t* wave_sums, uint64_t* total) {
    const int ln = static_cast<int>(threadIdx.x) & (kWave - 1);
    const int w = static_cast<int>(threadIdx.x) / kWave;
    unsigned long long inc = x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const unsigned long long y = __shfl_up(inc, o, kWave);
        if (ln >= o) {
            inc += y;
        }
    }
    if (ln == kWave - 1) {
        wave_sums[w] = inc;
    }
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kThreads / kWave; k++) {
        const uint64_t s = wave_sums[k];
        off += k < w ? s : 0;
        tot += s;
    }
    __syncthreads();  // wave_sums may be reused by the caller
    *total = tot;
    return off + inc - x;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                            uint64_t* __restrict__ part) {
    __shared__ uint64_t ws[kScanBlock / kWave];
    uint32_t v[kScanItems];
    load_counts(in, blockIdx.x * kScanTile + threadIdx.x * static_cast<uint64_t>(kScanItems), n, v);
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        s += v[k];
    }
    uint64_t tot = 0;
    (void)block_exclusive<kScanBlock>(s, ws, &tot);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tot;
    }
}

// One block: exclusive scan of the tile sums in place, and the grand total into out[n].
__global__ void __launch_bounds__(kTopBlock) k_scan_top(uint64_t* __restrict__ part, uint64_t tiles,
                                                        uint64_t* __restrict__ out_total) {
    __shared__ uint64_t ws[kTopBlock / kWave];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < tiles; base += static_cast<uint64_t>(kTopBlock) * kTopItems) {
        const uint64_t first = base + threadIdx.x * static_cast<uint64_t>(kTopItems);
        uint64_t v[kTopItems];
        uint64_t s = 0;
#pragma unroll
        for (int k = 0; k < kTopItems; k++) {
            v[k] = first + k < tiles ? part[first + k] : 0;
            s += v[k];
        }
        uint64_t tot = 0;
        uint64_t run = carry + block_exclusive<kTopBlock>(s, ws, &tot);
#pragma unroll
        for (int k = 0; k < kTopItems; k++) {
            if (first + k < tiles) {
                part[first + k] = run;
            }
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *out_total = carry;
    }
}

__global__ void __launch_bounds__(kScanBlock) k_scan_down(const uint32_t* __restrict__ in, uint64_t n,
                                                          const uint64_t* __restrict__ part, uint64_t* __restrict__ out) {
    __shared__ uint64_t ws[kScanBlock / kWave];
    const uint64_t first = blockIdx.x * kScanTile + threadIdx.x * static_cast<uint64_t>(kScanItems);
    uint32_t v[kScanItems];
    load_counts(in, first, n, v);
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        s += v[k];
    }
    uint64_t tot = 0;
    uint64_t run = part[blockIdx.x] + block_exclusive<kScanBlock>(s, ws, &tot);
    uint64_t o[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        o[k] = run;
        run += v[k];
    }
    if (first + kScanItems <= n) {
        ulonglong2* p = reinterpret_cast<ulonglong2*>(out + first);
#pragma unroll
        for (int k = 0; k < kScanItems / 2; k++) {
            p[k] = make_ulonglong2(o[2 * k], o[2 * k + 1]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            if (first + k < n) {
                out[first + k] = o[k];
            }
        }
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, uint64_t n, const uint64_t* __restrict__ n_dev,
                           uint64_t cap) {
    uint64_t count = n;
    if (n_dev != nullptr) {
        count = *n_dev < cap ? *n_dev : cap;
    }
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < count; i += stride) {
        p[i] = value;
    }
}

__global__ void k_set_u32(uint32_t* __restrict__ p, uint32_t a, uint32_t b) {
    if (threadIdx.x == 0) {
        p[0] = a;
        p[1] = b;
    }
}

}  // namespace

uint64_t scan_scratch_entries(uint64_t n) {
    return (n + kScanTile - 1) / kScanTile + 1;
}

hipError_t scan_counts_on(hipStream_t s, const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* part) {
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles > 0) {
        k_scan_reduce<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, part);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            return e;
        }
    }
    k_scan_top<<<1, kTopBlock, 0, s>>>(part, tiles, out + n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && tiles > 0) {
        k_scan_down<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, n, part, out);
        e = hipGetLastError();
    }
    return e;
}

int scan_counts(rl_bh_ctx* c, const uint32_t* in, uint64_t n, uint64_t* out) {
    if ((reinterpret_cast<uintptr_t>(in) & 15U) != 0 || (reinterpret_cast<uintptr_t>(out) & 15U) != 0) {
        return fail(c, RL_EINVAL, "scan_counts: operands must be 16-byte aligned");
    }
    if (scan_scratch_entries(n) > c->scan_part_cap) {
        return fail(c, RL_ESTATE, "scan_counts: scratch smaller than the scan");
    }
    RL_HIP(c, scan_counts_on(c->stream, in, n, out, c->scan_part));
    return RL_OK;
}

hipError_t fill_u32_on(hipStream_t s, uint32_t* p, uint32_t value, uint64_t n, const uint64_t* n_dev, uint64_t cap) {
    k_fill_u32<<<kStreamBlocks, 256, 0, s>>>(p, value, n, n_dev, cap);
    return hipGetLastError();
}

int fill_u32(rl_bh_ctx* c, uint32_t* p, uint32_t value, uint64_t n, const uint64_t* n_dev, uint64_t cap) {
    RL_HIP(c, fill_u32_on(c->stream, p, value, n, n_dev, cap));
    return RL_OK;
}

int set_steps(rl_bh_ctx* c, uint32_t neuron_step, uint32_t update) {
    k_set_u32<<<1, 64, 0, c->stream>>>(c->d_steps, neuron_step, update);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

}  // namespace rl
