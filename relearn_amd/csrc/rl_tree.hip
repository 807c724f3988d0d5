// Octree construction and the per-step upward pass on the device.
//
// Construction (once per rl_bh_init): every position gets a 126-bit key by replaying the octant
// decisions of the reference's insertion (Cell::get_octant_for_position, Cell.h:104-137: bit set when
// coordinate >= (min+max)/2; child box by midpoint, Cell.h:145-178) for 42 levels.  After a radix sort
// the capacity-1 octree is implied by the common prefix lengths of neighbouring keys:
//   c(p) = number of leading octant digits shared by sorted points p and p+1,
//   point p starts inner nodes at levels [A_p, c(p)] with A_p = c(p-1)+1 (A_0 = 0) and owns its
//   leaf at level max(c(p-1), c(p)) + 1.
// This is the tree OctreeNode::insert (OctreeNode.h:153-250) produces for distinct points in any
// insertion order: a node exists iff its box holds >= 2 points or it is a non-empty octant of such a
// node.  Nodes are sorted by (level, first point) = breadth-first with Morton order inside a level,
// so the children of a node are contiguous.
//
// Ensembles: every network gets its own key space (its own box); neurons are sorted by (network,
// key) and the common prefix between the last point of one network and the first of the next is
// taken as -1, so each network starts a root at level 0.  The forest is breadth-first over all
// networks together, so the roots are nodes 0..networks-1 and children stay contiguous.
//
// Upward pass (per plasticity step): leaves get the vacant dendrite counts
// (OctreeAlgorithm::update_leaf_nodes, OctreeAlgorithm.h:154-238), then one launch per level from the
// deepest to the root sums children in octant order and forms the count-weighted centroid in FP64
// exactly as OctreeNodeUpdater::update_node (OctreeNodeHelper.h:43-152): acc += pos * (double)n,
// pos = acc / (double)total, clamped to the node box.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "rl_device.h"
#include "rl_internal.h"

namespace rl {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}


__device__ __forceinline__ uint32_t digit_of(uint64_t khi, uint64_t klo, uint32_t level) {
    // level 1..21 in khi (level 1 most significant), 22..42 in klo
    if (level <= 21U) {
        return static_cast<uint32_t>((khi >> (3U * (21U - level))) & 7ULL);
    }
    return static_cast<uint32_t>((klo >> (3U * (42U - level))) & 7ULL);
}

// Key of every neuron by replaying the octant decisions level by level in its network's box.
__global__ void k_keys(uint64_t n, const double* __restrict__ pos, const uint32_t* __restrict__ net_of,
                       const double* __restrict__ net_box, uint64_t* __restrict__ khi, uint64_t* __restrict__ klo,
                       uint32_t* __restrict__ idx, int* __restrict__ err) {
#pragma clang fp contract(off)
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double p[3] = { pos[i], pos[n + i], pos[2 * n + i] };
    const double* box = net_box + 6 * static_cast<uint64_t>(net_of[i]);
    double lo[3] = { box[0], box[1], box[2] };
    double hi[3] = { box[3], box[4], box[5] };
    bool inside = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        inside = inside && (lo[d] <= p[d]) && (p[d] <= hi[d]);
    }
    if (!inside) {
        atomicOr(err, 1);
    }
    uint64_t a = 0, b = 0;
    for (uint32_t level = 1; level <= kMaxLevels; level++) {
        uint32_t dig = 0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double mid = (lo[d] + hi[d]) / 2.0;
            const bool upper = !(p[d] < mid);
            dig |= (upper ? 1U : 0U) << d;
            if (upper) {
                lo[d] = mid;
            } else {
                hi[d] = mid;
            }
        }
        if (level <= 21U) {
            a = (a << 3U) | dig;
        } else {
            b = (b << 3U) | dig;
        }
    }
    khi[i] = a;
    klo[i] = b;
    idx[i] = static_cast<uint32_t>(i);
}

__global__ void k_iota(uint64_t n, uint32_t* __restrict__ v) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        v[i] = static_cast<uint32_t>(i);
    }
}

__global__ void k_gather64(uint64_t n, const uint32_t* __restrict__ idx, const uint64_t* __restrict__ src,
                           uint64_t* __restrict__ dst) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        dst[i] = src[idx[i]];
    }
}

__global__ void k_any_equal(uint64_t n, const uint64_t* __restrict__ k, int* __restrict__ flag) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i + 1 < n && k[i] == k[i + 1]) {
        atomicOr(flag, 1);
    }
}

constexpr uint8_t kNoPrefix = 0xFF;   // lcp of two points of different networks (-1)

// c(p) for p in [0, n-1): shared leading digits of sorted keys p and p+1 (42 = identical keys), or
// kNoPrefix when p and p+1 belong to different networks.
__global__ void k_lcp(uint64_t n, const uint64_t* __restrict__ khi, const uint64_t* __restrict__ klo,
                      const uint32_t* __restrict__ netp, uint8_t* __restrict__ lcp, int* __restrict__ err) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i + 1 >= n) {
        return;
    }
    if (netp[i] != netp[i + 1]) {
        lcp[i] = kNoPrefix;
        return;
    }
    const uint64_t x = khi[i] ^ khi[i + 1];
    uint32_t c;
    if (x != 0) {
        c = (static_cast<uint32_t>(__clzll(static_cast<long long>(x))) - 1U) / 3U;
    } else {
        const uint64_t y = klo[i] ^ klo[i + 1];
        c = (y != 0) ? 21U + (static_cast<uint32_t>(__clzll(static_cast<long long>(y))) - 1U) / 3U : kMaxLevels;
    }
    if (c >= kMaxLevels) {
        atomicOr(err, 2);
    }
    lcp[i] = static_cast<uint8_t>(c);
}

__device__ __forceinline__ int lcp_at(const uint8_t* lcp, uint64_t n, int64_t p) {
    if (p < 0 || p + 1 >= static_cast<int64_t>(n) || lcp[p] == kNoPrefix) {
        return -1;
    }
    return static_cast<int>(lcp[p]);
}

// Network of every sorted position.
__global__ void k_net_of_sorted(uint64_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ net_of,
                                uint32_t* __restrict__ netp) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        netp[i] = net_of[order[i]];
    }
}

// out[i] = in[perm[i]] for 64-bit and 32-bit arrays
__global__ void k_permute64(uint64_t n, const uint32_t* __restrict__ perm, const uint64_t* __restrict__ in,
                            uint64_t* __restrict__ out) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        out[i] = in[perm[i]];
    }
}

__global__ void k_permute32(uint64_t n, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ in,
                            uint32_t* __restrict__ out) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        out[i] = in[perm[i]];
    }
}

__global__ void k_node_counts(uint64_t n, const uint8_t* __restrict__ lcp, uint64_t* __restrict__ cnt) {
    const uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
    const int cl = lcp_at(lcp, n, static_cast<int64_t>(p) - 1);
    const int cr = lcp_at(lcp, n, static_cast<int64_t>(p));
    const int a = cl + 1;
    const int leaf = max(cl, cr) + 1;
    cnt[p] = static_cast<uint64_t>(leaf - a + 1);
}

__global__ void k_emit_nodes(uint64_t n, const uint8_t* __restrict__ lcp, const uint64_t* __restrict__ off,
                             uint64_t* __restrict__ keys) {
    const uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (p >= n) {
        return;
    }
    const int cl = lcp_at(lcp, n, static_cast<int64_t>(p) - 1);
    const int cr = lcp_at(lcp, n, static_cast<int64_t>(p));
    const int a = cl + 1;
    const int leaf = max(cl, cr) + 1;
    uint64_t o = off[p];
    for (int l = a; l <= leaf; l++) {
        keys[o++] = (static_cast<uint64_t>(l) << 32U) | p;
    }
}

__global__ void k_level_starts(uint64_t m, const uint64_t* __restrict__ keys, uint32_t nlev, uint64_t* __restrict__ ls) {
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > nlev) {
        return;
    }
    // first index with level >= l
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((keys[mid] >> 32U) < l) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    ls[l] = lo;
}

// Per node: level, parent, box, max edge, octant; leaves get the neuron id and position.
__global__ void k_link(uint64_t m, uint64_t n, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ ls,
                       const uint8_t* __restrict__ lcp, const uint64_t* __restrict__ khi, const uint64_t* __restrict__ klo,
                       const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ netp,
                       const double* __restrict__ pos, const double* __restrict__ net_box,
                       uint8_t* __restrict__ level, uint8_t* __restrict__ octant, uint32_t* __restrict__ parent,
                       double* __restrict__ nbox, uint32_t* __restrict__ link, uint8_t* __restrict__ isleaf,
                       uint32_t* __restrict__ leaf_of, uint32_t* __restrict__ node_net, double4* __restrict__ rec0,
                       double4* __restrict__ rec1) {
#pragma clang fp contract(off)
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= m) {
        return;
    }
    const uint64_t key = keys[i];
    const uint32_t l = static_cast<uint32_t>(key >> 32U);
    const uint64_t p = key & 0xFFFFFFFFULL;
    level[i] = static_cast<uint8_t>(l);
    const uint64_t a = khi[p], b = klo[p];
    const uint32_t g = netp[p];
    node_net[i] = g;
    const double* box = net_box + 6 * static_cast<uint64_t>(g);
    double lo[3] = { box[0], box[1], box[2] };
    double hi[3] = { box[3], box[4], box[5] };
    for (uint32_t k = 1; k <= l; k++) {
        const uint32_t dig = digit_of(a, b, k);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double mid = (lo[d] + hi[d]) / 2.0;
            if ((dig >> d) & 1U) {
                lo[d] = mid;
            } else {
                hi[d] = mid;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        nbox[6 * i + d] = lo[d];
        nbox[6 * i + 3 + d] = hi[d];
    }
    // std::max({dx, dy, dz}): first maximum in argument order
    const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    double side = dx;
    if (side < dy) {
        side = dy;
    }
    if (side < dz) {
        side = dz;
    }
    octant[i] = l == 0 ? 0 : static_cast<uint8_t>(digit_of(a, b, l));
    const int cl = lcp_at(lcp, n, static_cast<int64_t>(p) - 1);
    const int cr = lcp_at(lcp, n, static_cast<int64_t>(p));
    const bool leaf = static_cast<int>(l) == max(cl, cr) + 1;
    isleaf[i] = leaf ? 1 : 0;
    if (leaf) {
        const uint32_t neuron = sorted[p];
        link[i] = neuron;
        leaf_of[neuron] = static_cast<uint32_t>(i);
        const double4 r = make_double4(pos[neuron], pos[n + neuron], pos[2 * n + neuron], side);
        rec0[i] = r;
        rec1[i] = r;
    } else {
        link[i] = 0;
        const double4 r = make_double4(0.0, 0.0, 0.0, side);
        rec0[i] = r;
        rec1[i] = r;
    }
    if (l == 0) {
        parent[i] = 0xFFFFFFFFU;
        return;
    }
    // parent: last node of level l-1 whose first point is <= p
    uint64_t lo_i = ls[l - 1], hi_i = ls[l];
    while (hi_i - lo_i > 1) {
        const uint64_t mid = (lo_i + hi_i) / 2;
        if ((keys[mid] & 0xFFFFFFFFULL) <= p) {
            lo_i = mid;
        } else {
            hi_i = mid;
        }
    }
    parent[i] = static_cast<uint32_t>(lo_i);
}

__global__ void k_children(uint64_t m, const uint64_t* __restrict__ ls, const uint8_t* __restrict__ level,
                           const uint32_t* __restrict__ parent, uint32_t* __restrict__ link,
                           uint32_t* __restrict__ nchild) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= m) {
        return;
    }
    const uint32_t l = level[i];
    if (l == 0) {
        return;
    }
    const uint32_t par = parent[i];
    if (i == ls[l] || parent[i - 1] != par) {
        link[par] = static_cast<uint32_t>(i);
    }
    atomicAdd(&nchild[par], 1U);
}

__global__ void k_pack_meta(uint64_t m, uint64_t n, const uint32_t* __restrict__ link,
                            const uint32_t* __restrict__ nchild, const uint8_t* __restrict__ isleaf,
                            uint2* __restrict__ meta0, uint2* __restrict__ meta1, int* __restrict__ err) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= m) {
        return;
    }
    const uint32_t nc = isleaf[i] ? 0U : nchild[i];
    // every index the traversal will derive from this node must lie inside the arrays
    const bool bad_inner = !isleaf[i] && (nc == 0 || nc > 8 || link[i] <= i || static_cast<uint64_t>(link[i]) + nc > m);
    const bool bad_leaf = isleaf[i] && static_cast<uint64_t>(link[i]) >= n;
    if (bad_inner || bad_leaf || link[i] > kLinkMask) {
        atomicOr(err, 4);
    }
    const uint2 v = make_uint2(0U, (link[i] & kLinkMask) | (nc << kLinkBits));
    meta0[i] = v;
    meta1[i] = v;
}

// ------------------------------------------------------------------ upward pass

__global__ void k_leaf_fill(uint64_t n, const uint32_t* __restrict__ leaf_of, const uint8_t* __restrict__ enabled,
                            const uint32_t* __restrict__ vexc, const uint32_t* __restrict__ vinh,
                            uint2* __restrict__ meta0, uint2* __restrict__ meta1) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t node = leaf_of[i];
    const bool on = enabled[i] != 0;
    meta0[node].x = on ? vexc[i] : 0U;
    meta1[node].x = on ? vinh[i] : 0U;
}

__device__ __forceinline__ uint32_t quantize(double v, double lo, double scale) {
#pragma clang fp contract(off)
    double q = rint((v - lo) * scale);
    q = q < 0.0 ? 0.0 : (q > 2147483647.0 ? 2147483647.0 : q);
    return static_cast<uint32_t>(q);
}

struct QuantArgs {
    const double* net_box;     // per network: lo xyz, hi xyz
    const double* net_scale;   // per network
    int enabled;
};

// Squared cell side in squared fixed-point units (the FP32 acceptance test compares it with the
// squared fixed-point distance).
__device__ __forceinline__ uint32_t quant_side2(double side, double scale) {
#pragma clang fp contract(off)
    const double sq = side * scale;
    return __float_as_uint(static_cast<float>(sq * sq));
}

// One thread per (node, dendrite type) of a level: the type's count and FP64 centroid of an inner node
// from its children (children loaded together, then summed in child order 0..7), clamped to the cell
// box, and its fixed-point record.
__global__ void k_upward_level(uint64_t first, uint64_t last, const double* __restrict__ nbox,
                               double4* __restrict__ rec0, double4* __restrict__ rec1, uint2* __restrict__ meta0,
                               uint2* __restrict__ meta1, uint4* __restrict__ q0, uint4* __restrict__ q1,
                               const uint32_t* __restrict__ node_net, QuantArgs qa) {
#pragma clang fp contract(off)
    const uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t i = first + (t >> 1U);
    const int s = static_cast<int>(t & 1U);
    if (i >= last) {
        return;
    }
    const uint32_t packed = meta0[i].y;
    const uint32_t nc = packed >> kLinkBits;
    if (nc == 0) {
        return;  // leaves keep the neuron position (Cell::set_neuron_position)
    }
    const uint32_t fc = packed & kLinkMask;
    double4* rec = s == 0 ? rec0 : rec1;
    uint2* meta = s == 0 ? meta0 : meta1;
    uint32_t cn[8];
    double4 cp[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        if (k < nc) {
            cn[k] = meta[fc + k].x;
            cp[k] = rec[fc + k];
        }
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    uint32_t total = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        if (k < nc) {
            total += cn[k];
            if (cn[k] != 0) {
                const double w = static_cast<double>(cn[k]);
                ax += cp[k].x * w;
                ay += cp[k].y * w;
                az += cp[k].z * w;
            }
        }
    }
    meta[i].x = total;
    double4 r = rec[i];
    if (total != 0) {
        const double w = static_cast<double>(total);
        double qx = ax / w, qy = ay / w, qz = az / w;
        const double* b = nbox + 6 * i;
        qx = qx < b[0] ? b[0] : (b[3] < qx ? b[3] : qx);
        qy = qy < b[1] ? b[1] : (b[4] < qy ? b[4] : qy);
        qz = qz < b[2] ? b[2] : (b[5] < qz ? b[5] : qz);
        r.x = qx;
        r.y = qy;
        r.z = qz;
    } else {
        r.x = 0.0;
        r.y = 0.0;
        r.z = 0.0;
    }
    rec[i] = r;
    if (qa.enabled) {
        uint4* q = s == 0 ? q0 : q1;
        const uint32_t g = node_net[i];
        const double* lo = qa.net_box + 6 * static_cast<uint64_t>(g);
        const double sc = qa.net_scale[g];
        q[i] = make_uint4(quantize(r.x, lo[0], sc), quantize(r.y, lo[1], sc), quantize(r.z, lo[2], sc),
                          quant_side2(r.w, sc));
    }
}

__global__ void k_quant_leaves(uint64_t m, const double4* __restrict__ rec, const uint32_t* __restrict__ node_net,
                               uint4* __restrict__ q0, uint4* __restrict__ q1, QuantArgs qa) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= m) {
        return;
    }
    const double4 r = rec[i];
    const uint32_t g = node_net[i];
    const double* lo = qa.net_box + 6 * static_cast<uint64_t>(g);
    const double sc = qa.net_scale[g];
    // a leaf's squared side is -infinity: theta^2 d^2 - side^2 is then +infinity, so the FP32 test accepts
    // the leaf, outside any guard band, as the acceptance criterion does for every leaf.  The upward
    // pass rewrites the records of inner nodes.
    const uint4 v = make_uint4(quantize(r.x, lo[0], sc), quantize(r.y, lo[1], sc), quantize(r.z, lo[2], sc),
                               0xFF800000U);
    q0[i] = v;
    q1[i] = v;
}

__global__ void k_quant_positions(uint64_t n, const double* __restrict__ pos, const uint32_t* __restrict__ net_of,
                                  uint32_t* __restrict__ qpos, QuantArgs qa) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t g = net_of[i];
    const double* lo = qa.net_box + 6 * static_cast<uint64_t>(g);
    const double sc = qa.net_scale[g];
    for (int d = 0; d < 3; d++) {
        qpos[d * n + i] = quantize(pos[d * n + i], lo[d], sc);
    }
}

QuantArgs quant_args(const rl_bh_ctx* c) {
    QuantArgs q{};
    q.net_box = c->net_box;
    q.net_scale = c->net_scale;
    q.enabled = c->cfg.precision == RL_PREC_FP32_MIXED ? 1 : 0;
    return q;
}

}  // namespace

void free_tree(rl_bh_ctx* c) {
    Tree& t = c->tree;
    dfree(&t.level);
    dfree(&t.octant);
    dfree(&t.box);
    for (int s = 0; s < 2; s++) {
        dfree(&t.rec64[s]);
        dfree(&t.rec32[s]);
        dfree(&t.wrec[s]);
        dfree(&t.meta[s]);
    }
    dfree(&t.node_net);
    t.m = 0;
    t.depth = 0;
    t.level_start.clear();
}

int build_tree(rl_bh_ctx* c, const double* h_xyz) {
    const uint64_t n = c->n;
    hipStream_t s = c->stream;
    free_tree(c);
    Tree& t = c->tree;
    if (n == 0) {
        return RL_OK;
    }
    // positions SoA on the device
    {
        std::vector<double> soa(3 * n);
        for (uint64_t i = 0; i < n; i++) {
            soa[i] = h_xyz[3 * i];
            soa[n + i] = h_xyz[3 * i + 1];
            soa[2 * n + i] = h_xyz[3 * i + 2];
        }
        RL_TRY(dalloc(c, &c->pos, 3 * n));
        RL_HIP(c, hipMemcpyAsync(c->pos, soa.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        RL_HIP(c, hipStreamSynchronize(s));
    }
    const uint32_t G = c->num_networks;
    uint64_t *khi = nullptr, *klo = nullptr, *khi2 = nullptr, *klo2 = nullptr;
    uint32_t *idx = nullptr, *idx2 = nullptr, *netp = nullptr, *netp2 = nullptr, *perm = nullptr;
    int* err = nullptr;
    uint8_t* lcp = nullptr;
    uint64_t* cnt = nullptr;
    uint64_t* off = nullptr;
    uint64_t *nkeys = nullptr, *nkeys2 = nullptr, *ls = nullptr;
    uint32_t *parent = nullptr, *link = nullptr, *nchild = nullptr;
    uint8_t* isleaf = nullptr;
    auto cleanup = [&]() {
        dfree(&khi);
        dfree(&klo);
        dfree(&khi2);
        dfree(&klo2);
        dfree(&idx);
        dfree(&idx2);
        dfree(&netp);
        dfree(&netp2);
        dfree(&perm);
        dfree(&err);
        dfree(&lcp);
        dfree(&cnt);
        dfree(&off);
        dfree(&nkeys);
        dfree(&nkeys2);
        dfree(&ls);
        dfree(&parent);
        dfree(&link);
        dfree(&nchild);
        dfree(&isleaf);
    };
    int status = RL_OK;
    do {
        if ((status = dalloc(c, &khi, n)) || (status = dalloc(c, &klo, n)) || (status = dalloc(c, &khi2, n)) ||
            (status = dalloc(c, &klo2, n)) || (status = dalloc(c, &idx, n)) || (status = dalloc(c, &idx2, n)) ||
            (status = dalloc(c, &err, 2)) || (status = dalloc(c, &lcp, n)) || (status = dalloc(c, &cnt, n + 1)) ||
            (status = dalloc(c, &off, n + 1)) || (status = dalloc(c, &netp, n)) || (status = dalloc(c, &netp2, n)) ||
            (status = dalloc(c, &perm, n))) {
            break;
        }
#define RL_STEP(expr)                                        \
    {                                                        \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) {                              \
            status = hip_fail(c, e_, #expr);                 \
            break;                                           \
        }                                                    \
    }
        RL_STEP(hipMemsetAsync(err, 0, 2 * sizeof(int), s));
        k_keys<<<grid_for(n), kBlock, 0, s>>>(n, c->pos, c->net_of, c->net_box, khi, klo, idx, err);
        RL_STEP(hipGetLastError());
        // sort by the first 21 levels; a second LSD pass over levels 22..42 only when needed
        size_t tb = 0;
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, khi, khi2, idx, idx2, static_cast<int>(n), 0, 63, s));
        if ((status = ensure_tmp(c, tb)) != RL_OK) {
            break;
        }
        RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tb, khi, khi2, idx, idx2, static_cast<int>(n), 0, 63, s));
        k_any_equal<<<grid_for(n), kBlock, 0, s>>>(n, khi2, err + 1);
        RL_STEP(hipGetLastError());
        int any_equal = 0;
        int herr = 0;
        RL_STEP(hipMemcpyAsync(&any_equal, err + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        RL_STEP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
        RL_STEP(hipStreamSynchronize(s));
        if (herr & 1) {
            status = fail(c, RL_EOUTSIDE, "a position lies outside the simulation box");
            break;
        }
        if (any_equal) {
            // full 126-bit order: stable LSD sort by the low key, then by the high key
            k_iota<<<grid_for(n), kBlock, 0, s>>>(n, idx);
            RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tb, klo, klo2, idx, idx2, static_cast<int>(n), 0, 63, s));
            k_gather64<<<grid_for(n), kBlock, 0, s>>>(n, idx2, khi, khi2);
            RL_STEP(hipGetLastError());
            RL_STEP(hipMemcpyAsync(khi, khi2, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
            RL_STEP(hipMemcpyAsync(idx, idx2, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tb, khi, khi2, idx, idx2, static_cast<int>(n), 0, 63, s));
        }
        // networks: a stable sort of the key order by network puts every network's points together,
        // each in key order
        k_net_of_sorted<<<grid_for(n), kBlock, 0, s>>>(n, idx2, c->net_of, netp);
        RL_STEP(hipGetLastError());
        if (G > 1) {
            int netbits = 1;
            while ((1ULL << netbits) < G) {
                netbits++;
            }
            // perm = sorted positions reordered stably by network (netp2 gets the sorted network ids)
            k_iota<<<grid_for(n), kBlock, 0, s>>>(n, idx);
            RL_STEP(hipGetLastError());
            size_t tbn = 0;
            RL_STEP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbn, netp, netp2, idx, perm, static_cast<int>(n), 0,
                                                       netbits, s));
            if ((status = ensure_tmp(c, tbn)) != RL_OK) {
                break;
            }
            RL_STEP(hipcub::DeviceRadixSort::SortPairs(c->tmp, tbn, netp, netp2, idx, perm, static_cast<int>(n), 0,
                                                       netbits, s));
            k_permute32<<<grid_for(n), kBlock, 0, s>>>(n, perm, idx2, idx);
            RL_STEP(hipGetLastError());
            k_permute64<<<grid_for(n), kBlock, 0, s>>>(n, perm, khi2, khi);
            RL_STEP(hipGetLastError());
            RL_STEP(hipMemcpyAsync(idx2, idx, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
            RL_STEP(hipMemcpyAsync(khi2, khi, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
            RL_STEP(hipMemcpyAsync(netp, netp2, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        }
        // idx2 = neurons in key order; khi2 = sorted high keys; gather the low keys
        k_gather64<<<grid_for(n), kBlock, 0, s>>>(n, idx2, klo, klo2);
        RL_STEP(hipGetLastError());
        if ((status = dalloc(c, &c->order, n)) != RL_OK) {
            break;
        }
        RL_STEP(hipMemcpyAsync(c->order, idx2, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        k_lcp<<<grid_for(n), kBlock, 0, s>>>(n, khi2, klo2, netp, lcp, err);
        RL_STEP(hipGetLastError());
        RL_STEP(hipMemsetAsync(cnt + n, 0, sizeof(uint64_t), s));
        k_node_counts<<<grid_for(n), kBlock, 0, s>>>(n, lcp, cnt);
        RL_STEP(hipGetLastError());
        {
            size_t tb2 = 0;
            RL_STEP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt, off, static_cast<int>(n) + 1, s));
            if ((status = ensure_tmp(c, tb2)) != RL_OK) {
                break;
            }
            RL_STEP(hipcub::DeviceScan::ExclusiveSum(c->tmp, tb2, cnt, off, static_cast<int>(n) + 1, s));
        }
        uint64_t m = 0;
        RL_STEP(hipMemcpyAsync(&m, off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        RL_STEP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
        RL_STEP(hipStreamSynchronize(s));
        if (herr & 2) {
            status = fail(c, RL_EDUPPOS, "two neurons share a position or lie closer than 42 octree levels resolve");
            break;
        }
        if (m > kLinkMask || n > kLinkMask) {
            status = fail(c, RL_ELIMIT, "more than 2^28 neurons or octree nodes");
            break;
        }
        if ((status = dalloc(c, &nkeys, m)) || (status = dalloc(c, &nkeys2, m))) {
            break;
        }
        k_emit_nodes<<<grid_for(n), kBlock, 0, s>>>(n, lcp, off, nkeys);
        RL_STEP(hipGetLastError());
        int endbit = 32;
        while ((1ULL << (endbit - 32)) <= kMaxLevels + 1) {
            endbit++;
        }
        size_t tb3 = 0;
        RL_STEP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb3, nkeys, nkeys2, static_cast<int>(m), 0, endbit, s));
        if ((status = ensure_tmp(c, tb3)) != RL_OK) {
            break;
        }
        RL_STEP(hipcub::DeviceRadixSort::SortKeys(c->tmp, tb3, nkeys, nkeys2, static_cast<int>(m), 0, endbit, s));
        const uint32_t nlev = kMaxLevels + 2;
        if ((status = dalloc(c, &ls, nlev + 1)) != RL_OK) {
            break;
        }
        k_level_starts<<<1, 64, 0, s>>>(m, nkeys2, nlev, ls);
        RL_STEP(hipGetLastError());
        std::vector<uint64_t> hls(nlev + 1);
        RL_STEP(hipMemcpyAsync(hls.data(), ls, (nlev + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        RL_STEP(hipStreamSynchronize(s));
        uint32_t depth = 0;
        for (uint32_t l = 0; l < nlev; l++) {
            if (hls[l + 1] > hls[l]) {
                depth = l;
            }
        }
        if (hls[1] != G) {
            status = fail(c, RL_EDUPPOS, "octree construction produced a forest without one root per network (internal error)");
            break;
        }
        t.m = m;
        t.depth = depth;
        t.level_start.assign(hls.begin(), hls.begin() + depth + 2);
        if ((status = dalloc(c, &t.level, m)) || (status = dalloc(c, &t.octant, m)) || (status = dalloc(c, &t.box, 6 * m)) ||
            (status = dalloc(c, &t.rec64[0], m)) || (status = dalloc(c, &t.rec64[1], m)) ||
            (status = dalloc(c, &t.meta[0], m)) || (status = dalloc(c, &t.meta[1], m)) ||
            (status = dalloc(c, &parent, m)) || (status = dalloc(c, &link, m)) || (status = dalloc(c, &nchild, m)) ||
            (status = dalloc(c, &isleaf, m)) || (status = dalloc(c, &c->leaf_of, n)) || (status = dalloc(c, &t.node_net, m))) {
            break;
        }
        if (c->cfg.precision == RL_PREC_FP32_MIXED) {
            if ((status = dalloc(c, &t.rec32[0], m)) || (status = dalloc(c, &t.rec32[1], m)) ||
                (status = dalloc(c, &t.wrec[0], 2 * (m + kNodePad))) || (status = dalloc(c, &t.wrec[1], 2 * (m + kNodePad))) ||
                (status = dalloc(c, &c->qpos, 3 * n))) {
                break;
            }
        }
        RL_STEP(hipMemsetAsync(nchild, 0, m * sizeof(uint32_t), s));
        k_link<<<grid_for(m), kBlock, 0, s>>>(m, n, nkeys2, ls, lcp, khi2, klo2, idx2, netp, c->pos, c->net_box, t.level,
                                               t.octant, parent, t.box, link, isleaf, c->leaf_of, t.node_net, t.rec64[0],
                                               t.rec64[1]);
        RL_STEP(hipGetLastError());
        k_children<<<grid_for(m), kBlock, 0, s>>>(m, ls, t.level, parent, link, nchild);
        RL_STEP(hipGetLastError());
        RL_STEP(hipMemsetAsync(err, 0, sizeof(int), s));
        k_pack_meta<<<grid_for(m), kBlock, 0, s>>>(m, n, link, nchild, isleaf, t.meta[0], t.meta[1], err);
        RL_STEP(hipGetLastError());
        if (c->cfg.precision == RL_PREC_FP32_MIXED) {
            const QuantArgs qa = quant_args(c);
            k_quant_leaves<<<grid_for(m), kBlock, 0, s>>>(m, t.rec64[0], t.node_net, t.rec32[0], t.rec32[1], qa);
            RL_STEP(hipGetLastError());
            k_quant_positions<<<grid_for(n), kBlock, 0, s>>>(n, c->pos, c->net_of, c->qpos, qa);
            RL_STEP(hipGetLastError());
        }
        RL_STEP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
        RL_STEP(hipStreamSynchronize(s));
        if (herr & 4) {
            status = fail(c, RL_EDUPPOS, "octree construction produced an invalid node (internal error)");
            break;
        }
#undef RL_STEP
    } while (false);
    cleanup();
    if (status != RL_OK) {
        free_tree(c);
    }
    return status;
}

int prepare_tree(rl_bh_ctx* c) {
    if (c->n == 0) {
        return RL_OK;
    }
    Tree& t = c->tree;
    hipStream_t s = c->stream;
    k_leaf_fill<<<grid_for(c->n), kBlock, 0, s>>>(c->n, c->leaf_of, c->enabled, c->vac[1], c->vac[2], t.meta[0],
                                                   t.meta[1]);
    RL_HIP(c, hipGetLastError());
    const QuantArgs qa = quant_args(c);
    for (int l = static_cast<int>(t.depth) - 1; l >= 0; l--) {
        const uint64_t first = t.level_start[l], last = t.level_start[l + 1];
        if (last <= first) {
            continue;
        }
        k_upward_level<<<grid_for(2 * (last - first)), kBlock, 0, s>>>(first, last, t.box, t.rec64[0], t.rec64[1],
                                                                        t.meta[0], t.meta[1], t.rec32[0], t.rec32[1],
                                                                        t.node_net, qa);
        RL_HIP(c, hipGetLastError());
    }
    return RL_OK;
}

int export_tree(rl_bh_ctx* c, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant, double* box_lo,
                double* box_hi, double* side, uint32_t* cnt_exc, uint32_t* cnt_inh, uint8_t* has_exc, uint8_t* has_inh,
                double* pos_exc, double* pos_inh) {
    Tree& t = c->tree;
    const uint64_t m = t.m;
    if (m == 0) {
        return RL_OK;
    }
    hipStream_t s = c->stream;
    std::vector<uint8_t> hlev(m), hoct(m);
    std::vector<double> hbox(6 * m);
    std::vector<double4> r0(m), r1(m);
    std::vector<uint2> m0(m), m1(m);
    RL_HIP(c, hipMemcpyAsync(hlev.data(), t.level, m, hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(hoct.data(), t.octant, m, hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(hbox.data(), t.box, 6 * m * sizeof(double), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(r0.data(), t.rec64[0], m * sizeof(double4), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(r1.data(), t.rec64[1], m * sizeof(double4), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(m0.data(), t.meta[0], m * sizeof(uint2), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(m1.data(), t.meta[1], m * sizeof(uint2), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipStreamSynchronize(s));
    for (uint64_t i = 0; i < m; i++) {
        const uint32_t nc = m0[i].y >> kLinkBits;
        const bool leaf = nc == 0;
        if (level) level[i] = hlev[i];
        if (octant) octant[i] = hoct[i];
        if (nchild) nchild[i] = static_cast<uint8_t>(nc);
        if (link) link[i] = m0[i].y & kLinkMask;
        if (box_lo) {
            for (int d = 0; d < 3; d++) box_lo[3 * i + d] = hbox[6 * i + d];
        }
        if (box_hi) {
            for (int d = 0; d < 3; d++) box_hi[3 * i + d] = hbox[6 * i + 3 + d];
        }
        if (side) side[i] = r0[i].w;
        if (cnt_exc) cnt_exc[i] = m0[i].x;
        if (cnt_inh) cnt_inh[i] = m1[i].x;
        const bool he = leaf || m0[i].x != 0;
        const bool hi = leaf || m1[i].x != 0;
        if (has_exc) has_exc[i] = he ? 1 : 0;
        if (has_inh) has_inh[i] = hi ? 1 : 0;
        if (pos_exc) {
            pos_exc[3 * i] = he ? r0[i].x : 0.0;
            pos_exc[3 * i + 1] = he ? r0[i].y : 0.0;
            pos_exc[3 * i + 2] = he ? r0[i].z : 0.0;
        }
        if (pos_inh) {
            pos_inh[3 * i] = hi ? r1[i].x : 0.0;
            pos_inh[3 * i + 1] = hi ? r1[i].y : 0.0;
            pos_inh[3 * i + 2] = hi ? r1[i].z : 0.0;
        }
    }
    return RL_OK;
}

}  // namespace rl
