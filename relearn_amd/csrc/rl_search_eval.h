// Acceptance test and probability of one octree node for one source, shared by the search kernels.
#pragma once

#include "rl_device.h"
#include "rl_internal.h"

namespace rl {

constexpr uint32_t kNone = 0xFFFFFFFFU;
constexpr uint32_t kOverflowMark = 0xFFFFFFFFU;   // found[src] of a source handed to the per-lane kernel

struct SearchArgs {
    uint64_t n;
    const uint64_t* nwork;  // device count of the sources in `order` for the per-lane kernel
    uint64_t m;          // tree nodes
    const uint32_t* order;
    const double* pos;
    const uint32_t* qpos;
    const uint8_t* signal;
    const uint8_t* enabled;
    const uint32_t* vac_axon;
    const uint64_t* slot_off;
    const double4* rec64[2];
    const uint4* rec32[2];
    const uint4* wrec[2];   // FP32 path: rec32 and meta side by side, two uint4 per node
    const uint2* meta[2];
    double theta;
    double scale64;      // -1 / sigma^2
    double mu;
    float theta2f;       // theta^2
    float k2f;           // -log2(e) / sigma^2 (micrometres)
    float muf;
    float guard;         // relative guard band of the FP32 acceptance test
    // absolute part of the band, in squared fixed-point units.  Rounding the source and the centroid
    // to the grid moves theta^2 d^2 by at most theta^2 (2 sqrt(3) d + 3) <= (guard/2)(|g| + s2) +
    // guard_abs/2, with g = theta^2 d^2 - s2 and guard_abs = 12 theta^2 / guard + 6 theta^2 (from
    // 2 sqrt(3) theta^2 d <= (guard/2) theta^2 d^2 + 6 theta^2 / guard), so outside
    // |g| <= guard s2 + guard_abs the FP32 sign of g is the exact one, for cells of any size.
    float guard_abs;
    // networks: network of each neuron, first neuron of each network, and per network the draw seed
    // and the fixed-point constants (-log2(e) / sigma^2 in squared units, micrometres per unit)
    const uint32_t* net_of;
    const uint64_t* net_off;
    const uint64_t* net_seed;
    const float* net_k2q;
    const float* net_invq;
    uint64_t seed;
    const uint32_t* step;  // device: plasticity update index (the draw-stream address)
    uint32_t* slot_target;
    uint32_t* slot_tests;
    uint32_t* found;
    uint32_t* searched;
    unsigned long long* counters;  // [0] executed tests, [1] reference tests, [2] searches
    uint32_t* redo;                // sources whose Gaussian sum underflowed (fallback launch)
    unsigned long long* n_redo;
};

// Per-source state; the record pointers are resolved once for the source's signal type so that the
// traversal loops index plain pointers instead of reloading them from the argument block.
struct Lane {
    double px, py, pz;
    int32_t qx, qy, qz;
    uint32_t src;        // neuron id in the context
    uint32_t local;      // neuron id inside its network (the draw-stream address)
    uint32_t root;       // root node of the network
    uint64_t seed;       // draw-stream seed of the network
    uint32_t step;       // plasticity update index
    float k2q, inv_q;    // fixed-point constants of the network
    int sig;
    const uint2* meta;
    const double4* rec64;
    const uint4* rec32;
};

__device__ __forceinline__ void lane_bind(const SearchArgs& a, Lane& L) {
    L.meta = L.sig ? a.meta[1] : a.meta[0];
    L.rec64 = L.sig ? a.rec64[1] : a.rec64[0];
    L.rec32 = L.sig ? a.rec32[1] : a.rec32[0];
}

// Everything a search needs about source neuron src (one lane's copy).
__device__ __forceinline__ void lane_source(const SearchArgs& a, Lane& L, uint32_t src, bool mixed) {
    L.src = src;
    L.sig = a.signal[src] == 0 ? 0 : 1;
    lane_bind(a, L);
    const uint32_t g = a.net_of[src];
    L.root = g;
    L.local = src - static_cast<uint32_t>(a.net_off[g]);
    L.seed = a.net_seed[g];
    L.step = *a.step;
    L.k2q = a.net_k2q[g];
    L.inv_q = a.net_invq[g];
    L.px = a.pos[src];
    L.py = a.pos[a.n + src];
    L.pz = a.pos[2 * a.n + src];
    if (mixed) {
        L.qx = static_cast<int32_t>(a.qpos[src]);
        L.qy = static_cast<int32_t>(a.qpos[a.n + src]);
        L.qz = static_cast<int32_t>(a.qpos[2 * a.n + src]);
    } else {
        L.qx = L.qy = L.qz = 0;
    }
}

enum : int { kDiscard = 0, kAccept = 1, kExpand = 2 };
enum : int { kGauss = 0, kInverse = 1 };

struct Eval {
    int status;
    bool exact;   // d64 valid (FP64 decision)
    double d64;
    float d2f;
};

// Exact FP64 acceptance test (BarnesHutBase.h:72-114) on a loaded record (centroid, max edge).
__device__ __forceinline__ Eval eval_exact_rec(const SearchArgs& a, const Lane& L, double4 r, bool leaf) {
    Eval e;
    e.exact = true;
    e.d2f = 0.0f;
    e.d64 = hypot3_exact(r.x - L.px, r.y - L.py, r.z - L.pz);
    if (e.d64 == 0.0) {
        e.status = kDiscard;
    } else if (leaf) {
        e.status = kAccept;
    } else {
        e.status = (r.w < (a.theta * e.d64)) ? kAccept : kExpand;
    }
    return e;
}

__device__ __forceinline__ Eval eval_exact(const SearchArgs& a, const Lane& L, uint32_t node, uint32_t cnt, bool leaf) {
    (void)cnt;
    return eval_exact_rec(a, L, L.rec64[node], leaf);
}

// FP32 test on the fixed-point record (x, y, z in fixed-point units, w = squared side in squared
// units): squared distance d2 = dx^2 + dy^2 + dz^2 of the exact integer differences, accept iff
// theta^2 d2 > side^2.  Decisions inside the guard band (and exact coincidences) are taken by the
// FP64 test.
__device__ __forceinline__ Eval eval_mixed_rec(const SearchArgs& a, const Lane& L, uint4 q, uint32_t node, bool leaf) {
    const int32_t ix = static_cast<int32_t>(q.x) - L.qx;
    const int32_t iy = static_cast<int32_t>(q.y) - L.qy;
    const int32_t iz = static_cast<int32_t>(q.z) - L.qz;
    if ((ix | iy | iz) == 0) {
        return eval_exact(a, L, node, 0, leaf);
    }
    const float dx = static_cast<float>(ix);
    const float dy = static_cast<float>(iy);
    const float dz = static_cast<float>(iz);
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    Eval e;
    e.exact = false;
    e.d64 = 0.0;
    e.d2f = d2;
    if (leaf) {
        e.status = kAccept;
        return e;
    }
    const float s2 = __uint_as_float(q.w);
    const float g = __builtin_fmaf(a.theta2f, d2, -s2);
    if (fabsf(g) <= __builtin_fmaf(a.guard, s2, a.guard_abs)) {
        return eval_exact(a, L, node, 0, leaf);
    }
    e.status = g > 0.0f ? kAccept : kExpand;
    return e;
}

__device__ __forceinline__ Eval eval_mixed(const SearchArgs& a, const Lane& L, uint32_t node, uint32_t cnt, bool leaf) {
    (void)cnt;
    return eval_mixed_rec(a, L, L.rec32[node], node, leaf);
}

// kMuZero: the caller guarantees mu == 0 (the default kernel), so the mixed path needs no square root.
template <bool kMixed, bool kMuZero = false>
__device__ __forceinline__ double probability(const SearchArgs& a, const Lane& L, const Eval& e, uint32_t cnt, int kind) {
#pragma clang fp contract(off)
    if (kind == kInverse) {
        return static_cast<double>(cnt) / e.d64;
    }
    if (!kMixed || e.exact) {
        // GaussianDistributionKernel::get_probability (source/algorithm/Kernel/Gaussian.h:62-67) times
        // the vacant count (Kernel.h:74)
        const double numerator = (e.d64 - a.mu) * (e.d64 - a.mu);
        const double exponent = numerator * a.scale64;
        return (exp(exponent) * 1.0) * static_cast<double>(cnt);
    }
    float x;
    if (kMuZero || a.muf == 0.0f) {
        x = e.d2f * L.k2q;
    } else {
        const float d = sqrtf(e.d2f) * L.inv_q - a.muf;
        x = d * d * a.k2f;
    }
    return static_cast<double>(exp2f(x)) * static_cast<double>(cnt);
}

}  // namespace rl
