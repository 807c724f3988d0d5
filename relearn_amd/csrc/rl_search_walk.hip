// Group walk: the FP32-mixed target search with one lane per source and one shared traversal.
//
// Active sources are grouped by (dendrite type searched, network) and, inside a group, taken 64 at a
// time in Morton order, so that the 64 sources of a wavefront lie close together.  Their level-1
// traversals (BarnesHutBase::get_nodes_to_consider, source/algorithm/BarnesHutInternal/
// BarnesHutBase.h:129-206) then visit nearly the same nodes: the wavefront walks the union of the 64
// traversals depth first, children in descending octant order (the reference's LIFO order), with one
// node per step and a lane mask per stack entry holding the sources that expanded the entry's node.
// Each lane evaluates the node only when it is in the mask (test_acceptance_criterion, :72-114, and
// the Gaussian attraction, Kernel.h:61-75, in FP32 with FP64 decisions at the acceptance threshold
// and at coincidences), so every lane sees exactly its own candidate list in the
// reference's order and adds the probabilities of its accepted nodes to a running FP64 sum in that
// order, as Kernel::create_probability_interval (Kernel.h:90-134) does.  The node data of a step is
// the same for all lanes (one broadcast load); the per-lane work is the test, the exponential and
// one FP64 addition.
//
// Every node that some lane expands gets a slot in the wave's table: the slots of its expanded
// children and, per lane that expanded it, the running sum when the walk leaves the node's subtree.
// A node none of whose children is expanded (most of them: their children are accepted or
// discarded by every lane) keeps no sums; its entry in the parent's slot says so, and a draw that
// meets it sums its children itself, in the walk's order.  A draw r of lane l is then resolved by
// lane l alone, top-down (ProbabilityPicker::pick_target, source/util/ProbabilityPicker.h:46-112):
// from the root it scans the children of its current node in list order, adds the probability of an
// accepted child, or jumps to the stored sum after an expanded child, and stops at the first accepted
// child whose running sum reaches r or enters the first expanded child whose stored sum does (or
// whose children, summed on, do).  The running sums are the same FP64 additions in the same order as
// the walk's, so the pick is the one of the sequential list.  Below a picked inner node the lane descends
// (BarnesHutBase::find_target_neuron, :223-255): the node's children, normally all accepted, are
// summed in list order, drawn from and picked; a level that needs expansion, or whose sum is 0 (the
// reference's fallbacks), is handed to the per-lane general traversal.
//
// Sources whose list has candidates but a sum of 0 (FP32 underflow; the FP64 and n/r fallbacks),
// sources with an FP64 acceptance comparison too close to call on squares, and groups whose table or
// stack would overflow go to the per-lane kernel whole.
//
// The walk pays off only when every wave has several groups (a group's walk is a chain of about 10^4
// dependent steps): an update with fewer active sources than walk_min_sources() takes the
// wave-per-source kernel, which is launched beside the walk; each returns at once unless the device's
// active-source count is on its side.
#include <algorithm>

#include "rl_search_eval.h"

namespace rl {
namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr uint32_t kGroup = 64;                  // sources per wavefront
constexpr uint32_t kWalkDepth = kMaxLevels + 4;  // stack entries (one per tree level)
constexpr int kXcds = kCursorSlots;
constexpr uint32_t kWalkCap = 3584;              // table slots per wave (expanded nodes of one group)
static_assert(kWalkCap < 0xFFFFU, "a slot and the no-slot mark fit the 16-bit child entries kept in LDS");
constexpr uint32_t kWalkWavesPerCu = 24;         // waves per CU the tables are sized for
constexpr uint32_t kNoGrandSlots = 0xFFFFFFFEU;  // child-slot entry of an expanded child none of whose children is expanded

inline unsigned grid_for(uint64_t n, int block = kBlock) {
    const uint64_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g == 0 ? 1 : g);
}

struct WalkArgs {
    SearchArgs s;
    const uint32_t* list;        // active sources by (type, network), Morton order inside
    const uint64_t* seg_goff;    // nseg + 1: first group of each segment (segment = type * G + network)
    const uint64_t* seg_first;   // nseg: list position of the segment's first source
    const uint64_t* seg_count;   // nseg: sources of the segment
    uint32_t nseg;
    uint32_t networks;
    unsigned long long* cursor;  // kXcds work cursors
    uint32_t chunk;              // groups per cursor increment
    double* t_cum;               // per wave: cap x 64 running sums at the exit of each slot's subtree
    uint32_t* t_child;           // per wave: cap x 8 slots of the expanded children (written for those only)
    uint32_t cap;
    uint32_t* overflow;          // sources for the per-lane kernel
    unsigned long long* n_overflow;
    uint4* slow;                 // (slot, source, node, level) of descents for the per-lane traversal
    unsigned long long* n_slow;
    const uint64_t* n_active;    // active sources of the update
    uint64_t run_from;           // the kernel returns at once when *n_active < run_from
};

constexpr uint32_t kBlockLevels = 18;            // LDS blocks of staged child records (a ring over the levels)

// The stack below the top entry, and the records of the children of the stack's nodes: each push
// copies the node's children from the tree into the block of its level (load-to-LDS instructions), so
// that the steps over a node's children read them from LDS instead of waiting for memory one child at
// a time.  Levels share blocks modulo kBlockLevels; a block records the level it holds, and a pop to
// a level whose block was taken by a deeper one copies the children again.
struct WalkShared {
    uint4 ent[kWalkDepth];   // (first child | next child to visit << 28, slot, lane mask low, high)
    uint16_t kids[kWalkDepth][8];   // entries of the children of each level's node, until it is left
    uint32_t owner[kBlockLevels];
    uint4 rec[kBlockLevels][16];   // per child: fixed-point record, then the count/link pair (low half)
};

__device__ __forceinline__ uint32_t block_of(uint32_t level) {
    return level % kBlockLevels;
}

__device__ __forceinline__ uint32_t lane_id() {
    return __lane_id();
}

__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32U),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0U));
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// The source data a lane needs for the FP32 test and, on the rare FP64 re-decision, for the exact one.
struct WalkLane {
    int32_t qx, qy, qz;
    uint32_t src;
};

struct Group {
    float k2q, inv_q;
    const uint4* wrec;   // per node: fixed-point record, then the count/link pair (low half)
    const double4* rec64;
};

__device__ __forceinline__ uint4 node_rec(const Group& G, uint32_t y) {
    return G.wrec[2 * static_cast<uint64_t>(y)];
}

__device__ __forceinline__ uint2 node_meta(const Group& G, uint32_t y) {
    return *reinterpret_cast<const uint2*>(G.wrec + 2 * static_cast<uint64_t>(y) + 1);
}

typedef __attribute__((address_space(3))) void LdsVoid;

// s_waitcnt immediate: VM_CNT 0 (the staged records have landed in LDS), EXP_CNT and LGKM_CNT not waited for
constexpr int kWaitVm0 = 0x0F70;

// Copies the records of the children [first, first + nc) of the node at stack level `level` into its
// LDS block (lanes < 2 nc a 16-byte half of a node's 32-byte record each) and waits for them.  A copy
// again after a deeper level took the block passes nc = 8 (the stack keeps no child count): records
// past the node's last child are then copied too and never read, and the node arrays end with
// kNodePad spare entries, so the copy stays inside them for every node.
__device__ __forceinline__ void stage_children(WalkShared& sh, const Group& G, uint32_t level, uint32_t first,
                                               uint32_t nc, uint32_t ln) {
    const uint32_t b = block_of(level);
    if (ln < 2U * nc) {
        __builtin_amdgcn_global_load_lds((void*)(G.wrec + 2 * static_cast<uint64_t>(first) + ln),
                                         (LdsVoid*)&sh.rec[b][0], 16, 0, 0);
    }
    if (__builtin_amdgcn_inverse_ballot_w64(1ULL)) {
        sh.owner[b] = level;
    }
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
}

struct NodeEval {
    int st;
    double p;
};

enum : int { kUndecided = 3 };

// The FP64 acceptance decision (test_acceptance_criterion, BarnesHutBase.h:72-114) of a node whose FP32
// test is not conclusive: discard at distance 0, accept a leaf; for an inner node, accept iff the side
// is below theta times the distance.  The comparison is made on squares (side^2 against theta^2 d^2,
// each within a few units in the last place of the exact values) with a relative margin of 1e-13
// around equality, outside of which it agrees with the reference's comparison of the side against
// theta times the libstdc++ hypot of the centroid difference; inside the margin it is kUndecided and
// the source goes to the per-lane search, which makes the reference's comparison itself.
__device__ __forceinline__ int decide_exact(const SearchArgs& a, const Group& G, uint32_t src, uint32_t node,
                                            bool leaf) {
#pragma clang fp contract(off)
    const double4 r = G.rec64[node];
    const double dx = r.x - a.pos[src];
    const double dy = r.y - a.pos[a.n + src];
    const double dz = r.z - a.pos[2 * a.n + src];
    if (dx == 0.0 && dy == 0.0 && dz == 0.0) {
        return kDiscard;
    }
    if (leaf) {
        return kAccept;
    }
    const double t = (a.theta * a.theta) * (dx * dx + dy * dy + dz * dz);
    const double s2 = r.w * r.w;
    if (s2 < t * (1.0 - 1.0e-13)) {
        return kAccept;
    }
    if (s2 > t * (1.0 + 1.0e-13)) {
        return kExpand;
    }
    return kUndecided;
}

// Acceptance test and probability of a loaded node for lanes with `act`: the FP32 test on exact
// fixed-point differences, the Gaussian as exp2 of the scaled squared distance times the vacant count
// in FP32.  Decisions inside the guard band and fixed-point coincidences are re-decided in FP64, so the
// candidate lists are the FP64 path's; the probabilities stay FP32.  Other lanes get kDiscard, p = 0.
template <bool kMu0>
__device__ __forceinline__ NodeEval eval_walk(const SearchArgs& a, const Group& G, const WalkLane& W, uint32_t node,
                                              uint2 mt, uint4 rq, bool act) {
    const bool valid = act && mt.x != 0;
    const bool leaf = (mt.y >> kLinkBits) == 0;
    const int32_t ix = static_cast<int32_t>(rq.x) - W.qx;
    const int32_t iy = static_cast<int32_t>(rq.y) - W.qy;
    const int32_t iz = static_cast<int32_t>(rq.z) - W.qz;
    const float dx = static_cast<float>(ix);
    const float dy = static_cast<float>(iy);
    const float dz = static_cast<float>(iz);
    const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    const float s2 = __uint_as_float(rq.w);
    const float g = __builtin_fmaf(a.theta2f, d2, -s2);
    const bool coinc = (ix | iy | iz) == 0;
    const bool near = !leaf && fabsf(g) <= __builtin_fmaf(a.guard, s2, a.guard_abs);
    float x;
    if (kMu0) {
        x = d2 * G.k2q;
    } else {
        const float dd = sqrtf(d2) * G.inv_q - a.muf;
        x = dd * dd * a.k2f;
    }
    const float w = __builtin_amdgcn_exp2f(x) * static_cast<float>(mt.x);
    NodeEval e;
    e.st = valid ? ((leaf || g > 0.0f) ? kAccept : kExpand) : kDiscard;
    e.p = e.st == kAccept ? static_cast<double>(w) : 0.0;
    // the source's own leaf lies at distance 0: discarded without the FP64 test
    const bool self = leaf && (mt.y & kLinkMask) == W.src;
    const bool need = valid && !self && (coinc || near);
    if (valid && self) {
        e.st = kDiscard;
        e.p = 0.0;
    }
    if (__builtin_amdgcn_ballot_w64(need) != 0) {
        if (need) {
            e.st = decide_exact(a, G, W.src, node, leaf);
            e.p = e.st == kAccept ? static_cast<double>(w) : 0.0;
        }
    }
    return e;
}

// Picker condition of ProbabilityPicker::pick_target: the first candidate whose running sum reaches r,
// or, when r is not positive, the first candidate with a positive probability.
__device__ __forceinline__ bool reaches(double sum, double r) {
    return r > 0.0 ? sum >= r : sum > 0.0;
}

enum : int { kDescDone = 0, kDescFailed = 1, kDescSlow = 2 };

// Descent below an accepted inner node of the lane's own list, level by level (find_target_neuron
// repeated on the picked node): the children are tested, summed in list order (descending octant),
// drawn from and picked.  A level with an expanded child or with a sum of 0 goes to the per-lane
// general traversal (kDescSlow with node and level).  tests gets the reference's tests of the levels.
template <bool kMu0>
__device__ int descend(const SearchArgs& a, const Group& G, const WalkLane& W, uint64_t seed, uint32_t local,
                       uint32_t step, uint32_t j, uint32_t& node, uint32_t& level, uint32_t& tests, uint32_t& exec) {
    for (uint32_t guard = 0; guard < kMaxLevels + 2; guard++) {
        if (node >= a.m) {
            return kDescFailed;
        }
        const uint2 mn = node_meta(G, node);
        const uint32_t nc = mn.y >> kLinkBits;
        if (nc == 0) {
            node = mn.y & kLinkMask;  // leaf: the target neuron
            return kDescDone;
        }
        if (mn.x == 0) {
            return kDescFailed;
        }
        const uint32_t first = mn.y & kLinkMask;
        // the children in list order (descending octant): statuses and the sum
        double tot = 0.0;
        bool expand = false;
        uint32_t accepted = 0;
        for (uint32_t k = nc; k-- > 0;) {
            const uint32_t y = first + k;
            const NodeEval e = eval_walk<kMu0>(a, G, W, y, node_meta(G, y), node_rec(G, y), true);
            expand = expand || e.st == kExpand || e.st == kUndecided;
            if (e.st == kAccept) {
                tot += e.p;
                accepted++;
            }
        }
        exec += nc;
        if (expand) {
            return kDescSlow;
        }
        tests += nc;
        if (accepted == 0) {
            return kDescFailed;
        }
        if (tot == 0.0) {
            return kDescSlow;
        }
        const double r = draw_to_uniform(pick_draw(seed, local, step, j, level), nextafter(tot, tot + 0.00001));
        // the same children again: the first whose running sum reaches r
        double cum = 0.0;
        int sel = -1;
        for (uint32_t k = nc; k-- > 0;) {
            const uint32_t y = first + k;
            const NodeEval e = eval_walk<kMu0>(a, G, W, y, node_meta(G, y), node_rec(G, y), true);
            exec++;
            if (e.st == kAccept) {
                const double cn = cum + e.p;
                if (reaches(cn, r)) {
                    sel = static_cast<int>(k);
                    break;
                }
                cum = cn;
            }
        }
        if (sel < 0) {
            return kDescSlow;
        }
        node = first + static_cast<uint32_t>(sel);
        level++;
    }
    return kDescFailed;
}

// The walk's node records, packed for an update that the walk searches: per node and dendrite type the
// fixed-point record and the count/link pair side by side (32 B), so that one load brings both (a push
// copies a node's children in one instruction, a draw's read of a node touches one cache line).  One
// thread per node, both types; returns at once when the wave-per-source kernel takes the update.
__global__ void k_pack_walk(uint64_t m, const uint64_t* __restrict__ n_active, uint64_t run_from,
                            const uint2* __restrict__ meta0, const uint2* __restrict__ meta1,
                            const uint4* __restrict__ q0, const uint4* __restrict__ q1, uint4* __restrict__ w0,
                            uint4* __restrict__ w1) {
    if (*n_active < run_from) {
        return;
    }
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= m) {
        return;
    }
    const uint2 a = meta0[i];
    const uint2 b = meta1[i];
    w0[2 * i] = q0[i];
    w0[2 * i + 1] = make_uint4(a.x, a.y, 0U, 0U);
    w1[2 * i] = q1[i];
    w1[2 * i + 1] = make_uint4(b.x, b.y, 0U, 0U);
}

template <bool kMu0, bool kCount>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(6, 6))) k_walk(WalkArgs a) {
    __shared__ WalkShared sh;
    // kCount: child records the wave examined (children of the roots and of expanded nodes), table slots used
    __shared__ unsigned long long cnt_steps, cnt_slots, cnt_bytes;
    if (*a.n_active < a.run_from) {
        return;   // the wave-per-source kernel takes this update
    }
    const SearchArgs& s = a.s;
    const uint32_t ln = lane_id();
    const uint64_t cbase = static_cast<uint64_t>(blockIdx.x) * a.cap;   // first table slot of this wave
    double* const tcum = a.t_cum + cbase * kWave;
    uint32_t* const tchild = a.t_child + cbase * 8;
    int seg_x = static_cast<int>(blockIdx.x % kXcds);
    int tried = 0;
    unsigned long long acc_exec = 0;
    const uint64_t nw = a.seg_goff[a.nseg];
    const uint32_t step = *s.step;
    const uint32_t chunk = a.chunk;
    if (kCount && ln == 0) {
        cnt_steps = 0;
        cnt_slots = 0;
        cnt_bytes = 0;
    }
    uint64_t w_next = 0, w_lim = 0;
    while (true) {
        if (w_next >= w_lim) {
            const uint64_t seg_end = nw * static_cast<uint64_t>(seg_x + 1) / kXcds;
            uint64_t w0 = 0;
            if (ln == 0) {
                w0 = nw * static_cast<uint64_t>(seg_x) / kXcds +
                     atomicAdd(&a.cursor[seg_x], static_cast<unsigned long long>(chunk));
            }
            w0 = (static_cast<uint64_t>(uniform_u32(static_cast<uint32_t>(w0 >> 32U))) << 32U) |
                 uniform_u32(static_cast<uint32_t>(w0));
            if (w0 >= seg_end) {
                if (++tried >= kXcds) {
                    break;
                }
                seg_x = (seg_x + 1) % kXcds;
                continue;
            }
            w_next = w0;
            w_lim = std::min<uint64_t>(w0 + chunk, seg_end);
        }
        const uint64_t w = w_next++;
        // segment of group w: the last segment whose first group is <= w
        uint32_t lo = 0, hi = a.nseg;   // seg_goff[lo] <= w < seg_goff[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (a.seg_goff[mid] <= w) {
                lo = mid;
            } else {
                hi = mid;
            }
        }
        const uint32_t seg = uniform_u32(lo);
        const uint64_t gi = w - a.seg_goff[seg];
        const uint32_t type = seg / a.networks;
        const uint32_t net = seg - type * a.networks;
        const uint64_t pos0 = a.seg_first[seg] + kGroup * gi;
        const uint64_t left = a.seg_count[seg] - kGroup * gi;
        const uint32_t gsize = static_cast<uint32_t>(std::min<uint64_t>(left, kGroup));
        Group G;
        G.wrec = s.wrec[type];
        G.rec64 = s.rec64[type];
        G.k2q = s.net_k2q[net];
        G.inv_q = s.net_invq[net];
        const uint64_t seed = s.net_seed[net];
        const uint32_t net_first = static_cast<uint32_t>(s.net_off[net]);
        const bool active = ln < gsize;
        WalkLane W;
        W.src = active ? a.list[pos0 + ln] : 0U;
        W.qx = static_cast<int32_t>(s.qpos[W.src]);
        W.qy = static_cast<int32_t>(s.qpos[s.n + W.src]);
        W.qz = static_cast<int32_t>(s.qpos[2 * s.n + W.src]);
        const uint64_t amask = __builtin_amdgcn_ballot_w64(active);

        // ---- the union walk: level-1 sums of all lanes ----
        // wave-uniform: the root values seed the traversal state, which must stay in scalar registers
        const uint2 mr0 = node_meta(G, net);
        const uint2 mr = make_uint2(uniform_u32(mr0.x), uniform_u32(mr0.y));
        double cum = 0.0;
        uint32_t tests = 0;
        uint64_t has = 0;
        uint64_t handoff = 0;   // sources with an FP64 decision too close to call here
        bool over = false;
        const uint32_t root_first = mr.y & kLinkMask;
        const uint32_t root_nc = mr.x == 0 ? 0U : (mr.y >> kLinkBits);
        if (root_nc != 0) {
            uint32_t top_first = root_first, top_cur = root_nc, top_slot = 0;
            uint64_t top_mask = amask;
            uint32_t nslots = 1, sp = 0;
            if (ln < kBlockLevels) {
                sh.owner[ln] = kNone;
            }
            stage_children(sh, G, 0, root_first, root_nc, ln);
            // a lane's level-1 tests: the root's children and the children of every node it expands
            tests = active ? root_nc : 0U;
            if (kCount && ln == 0) {
                atomicAdd(&cnt_steps, static_cast<unsigned long long>(root_nc));
            }
            while (true) {
                // steps over the remaining children of the top entry, until some lane expands one or
                // none is left; only the child counter and the sums change here
                uint64_t exm = 0;
                {
                    const uint32_t bl = block_of(sp);
                    while (exm == 0 && top_cur != 0) {
                        top_cur--;
                        const uint2 mraw = *reinterpret_cast<const uint2*>(&sh.rec[bl][2 * top_cur + 1]);
                        const uint4 rq = sh.rec[bl][2 * top_cur];
                        // both records are read together, before the count decides anything
                        __asm__ volatile("" ::"v"(mraw.x), "v"(mraw.y), "v"(rq.x), "v"(rq.y), "v"(rq.z), "v"(rq.w));
                        const uint32_t cnt = uniform_u32(mraw.x);
                        if (cnt == 0) {
                            continue;   // no vacant dendrites below: discarded by every lane
                        }
                        const int32_t ix = static_cast<int32_t>(rq.x) - W.qx;
                        const int32_t iy = static_cast<int32_t>(rq.y) - W.qy;
                        const int32_t iz = static_cast<int32_t>(rq.z) - W.qz;
                        const float dx = static_cast<float>(ix);
                        const float dy = static_cast<float>(iy);
                        const float dz = static_cast<float>(iz);
                        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        // accepted when theta^2 d^2 > side^2 clear of the guard band, expanded when below
                        // it.  A leaf's squared side is -infinity, so a leaf is accepted, never near the
                        // band.  The squared distance of integer differences is 0 exactly at a
                        // fixed-point coincidence, which is decided in FP64 (the source's own leaf, a
                        // coincidence since a leaf keeps its neuron's position, is dropped there).  Lane
                        // masks come straight from the comparisons and are combined with scalar mask
                        // arithmetic.
                        const float s2 = __uint_as_float(rq.w);
                        const float g = __builtin_fmaf(s.theta2f, d2, -s2);
                        const uint64_t coinc = __builtin_amdgcn_ballot_w64(d2 == 0.0f);
                        const uint64_t near =
                            __builtin_amdgcn_ballot_w64(fabsf(g) <= __builtin_fmaf(s.guard, s2, s.guard_abs));
                        const uint64_t pos = __builtin_amdgcn_ballot_w64(g > 0.0f);
                        float x;
                        if (kMu0) {
                            x = d2 * G.k2q;
                        } else {
                            const float dd = sqrtf(d2) * G.inv_q - s.muf;
                            x = dd * dd * s.k2f;
                        }
                        const float w = __builtin_amdgcn_exp2f(x) * static_cast<float>(cnt);
                        const uint64_t unsure = coinc | near;
                        const uint64_t sure = top_mask & ~unsure;
                        const uint64_t need = top_mask & unsure;
                        uint64_t acc = sure & pos;
                        uint64_t ex = sure & ~pos;
                        if (need != 0) {
                            const uint32_t mt_y = uniform_u32(mraw.y);
                            const bool leaf = (mt_y >> kLinkBits) == 0;
                            int st = kDiscard;
                            if (__builtin_amdgcn_inverse_ballot_w64(need)) {
                                if (!leaf || (mt_y & kLinkMask) != W.src) {
                                    st = decide_exact(s, G, W.src, top_first + top_cur, leaf);
                                }
                            }
                            acc |= __builtin_amdgcn_ballot_w64(st == kAccept);
                            ex |= __builtin_amdgcn_ballot_w64(st == kExpand);
                            handoff |= __builtin_amdgcn_ballot_w64(st == kUndecided);
                        }
                        float wa = __builtin_amdgcn_inverse_ballot_w64(acc) ? w : 0.0f;
                        __asm__ volatile("" : "+v"(wa));   // select in FP32, then one conversion
                        cum += static_cast<double>(wa);
                        has |= acc;
                        exm = ex;
                    }
                }
                if (exm != 0) {
                    // some lanes expand child top_cur: its slot and children become the top entry
                    if (nslots >= a.cap || sp + 1 >= kWalkDepth) {
                        over = true;
                        break;
                    }
                    const uint32_t ns = nslots;
                    nslots = uniform_u32(ns + 1);
                    const uint32_t link = uniform_u32(sh.rec[block_of(sp)][2 * top_cur + 1].y);
                    if (__builtin_amdgcn_inverse_ballot_w64(1ULL)) {
                        sh.ent[sp] = make_uint4(top_first | (top_cur << kLinkBits), top_slot,
                                                static_cast<uint32_t>(top_mask), static_cast<uint32_t>(top_mask >> 32U));
                    }
                    sp++;
                    top_first = link & kLinkMask;
                    top_cur = link >> kLinkBits;
                    top_slot = ns;
                    top_mask = exm;
                    if (__builtin_amdgcn_inverse_ballot_w64(exm)) {
                        tests += top_cur;
                    }
                    if (kCount && ln == 0) {
                        atomicAdd(&cnt_steps, static_cast<unsigned long long>(top_cur));
                    }
                    stage_children(sh, G, sp, top_first, top_cur, ln);
                    continue;
                }
                // the walk leaves the subtree of this slot's node
                if (sp == 0) {
                    // the root's sums are the lanes' totals, which the draws take from registers; its
                    // children's entries are stored
                    if (__builtin_amdgcn_inverse_ballot_w64(0xFFULL)) {
                        const uint32_t v = sh.kids[0][ln];
                        tchild[ln] = v == 0xFFFFU ? kNoGrandSlots : v;
                    }
                    if (kCount && ln == 0) {
                        atomicAdd(&cnt_bytes, 32ULL);
                    }
                    if (kCount && ln == 0) {
                        atomicAdd(&cnt_slots, static_cast<unsigned long long>(nslots));
                    }
                    break;
                }
                // the running sums of the lanes that expanded the node, the only lanes that read them
                // back, unless no child of the node was expanded: a draw then sums the node's children
                // itself (as many as a table entry would save it), in the walk's order
                const uint32_t cslot = top_slot;
                const bool grand = nslots != cslot + 1U;
                if (grand) {
                    // the table is written once and read back at most once, by the draws: its stores
                    // bypass the caches' retention (non-temporal), which keeps L2 for the node records
                    if (__builtin_amdgcn_inverse_ballot_w64(top_mask)) {
                        __builtin_nontemporal_store(cum, &tcum[static_cast<uint64_t>(cslot) * kWave + ln]);
                    }
                    // the entries of the node's children, in one store: written for expanded children
                    // only, and a lane reads only those of the children it expanded
                    if (__builtin_amdgcn_inverse_ballot_w64(0xFFULL)) {
                        const uint32_t v = sh.kids[sp][ln];
                        __builtin_nontemporal_store(v == 0xFFFFU ? kNoGrandSlots : v, &tchild[static_cast<uint64_t>(cslot) * 8 + ln]);
                    }
                    if (kCount && ln == 0) {
                        atomicAdd(&cnt_bytes, static_cast<unsigned long long>(8 * __popcll(top_mask) + 32));
                    }
                }
                sp--;
                const uint4 e = sh.ent[sp];
                const uint32_t e0 = uniform_u32(e.x);
                top_first = e0 & kLinkMask;
                top_cur = e0 >> kLinkBits;
                top_slot = uniform_u32(e.y);
                top_mask = (static_cast<uint64_t>(uniform_u32(e.w)) << 32U) | uniform_u32(e.z);
                // the child's entry in the parent's slot
                if (__builtin_amdgcn_inverse_ballot_w64(1ULL)) {
                    sh.kids[sp][top_cur] = static_cast<uint16_t>(grand ? cslot : 0xFFFFU);
                }
                if (uniform_u32(sh.owner[block_of(sp)]) != sp) {
                    stage_children(sh, G, sp, top_first, 8U, ln);
                }
            }
        }
        // the table written by other lanes is read below
        __threadfence_block();
        const uint32_t na = active ? s.vac_axon[W.src] : 0U;
        const uint64_t slot0 = active ? s.slot_off[W.src] : 0U;
        const double total = cum;
        const bool mine = ((has >> ln) & 1ULL) != 0;
        acc_exec += tests;
        // sources for the per-lane kernel: the whole group on an overflow, else lists summing to 0
        const bool to_lane = active && (over || (mine && total == 0.0) || ((handoff >> ln) & 1ULL) != 0);
        const uint64_t om = __builtin_amdgcn_ballot_w64(to_lane);
        if (om != 0) {
            unsigned long long o0 = 0;
            if (ln == 0) {
                o0 = atomicAdd(a.n_overflow, static_cast<unsigned long long>(__popcll(om)));
            }
            o0 = __shfl(o0, 0, kWave);
            if (to_lane) {
                a.overflow[o0 + rank_below(om)] = W.src;
                s.found[W.src] = kOverflowMark;
            }
        }
        if (active && !to_lane && !mine) {
            s.slot_tests[slot0] = tests;   // an empty list: the first search fails, slot_target stays kNone
        }
        const bool draws = active && !to_lane && mine;
        const double next1 = draws ? nextafter(total, total + 0.00001) : 0.0;
        // the draws of all lanes, numbered lane by lane (a lane's draws j = 0 .. na - 1 in order), are
        // shared out over the wave 64 at a time: item i goes to lane i mod 64, which resolves it for
        // the source lane holding it, with that lane's position, sums and table column
        const uint32_t nd = draws ? na : 0U;
        uint32_t incl = nd;   // inclusive prefix sum of the draws over the lanes
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o, kWave));
            incl += ln >= static_cast<uint32_t>(o) ? t : 0U;
        }
        const uint32_t ndraws = uniform_u32(static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWave - 1, kWave)));
        uint32_t exec = 0;
        for (uint32_t base = 0; base < ndraws; base += kWave) {
            const uint32_t item = base + ln;
            // the source lane of the item: the first lane whose inclusive prefix exceeds it
            uint32_t sl = 0;
#pragma unroll
            for (uint32_t half = kWave / 2; half > 0; half >>= 1) {
                const uint32_t probe = sl + half - 1U;
                const uint32_t at = static_cast<uint32_t>(__shfl(static_cast<int>(incl), static_cast<int>(probe), kWave));
                sl = at <= item ? sl + half : sl;
            }
            bool nav = item < ndraws;
            sl = nav ? sl : ln;
            const uint32_t j = item - (static_cast<uint32_t>(__shfl(static_cast<int>(incl), static_cast<int>(sl), kWave)) -
                                       static_cast<uint32_t>(__shfl(static_cast<int>(nd), static_cast<int>(sl), kWave)));
            WalkLane V;
            V.qx = __shfl(W.qx, static_cast<int>(sl), kWave);
            V.qy = __shfl(W.qy, static_cast<int>(sl), kWave);
            V.qz = __shfl(W.qz, static_cast<int>(sl), kWave);
            V.src = static_cast<uint32_t>(__shfl(static_cast<int>(W.src), static_cast<int>(sl), kWave));
            const double vnext1 = __shfl(next1, static_cast<int>(sl), kWave);
            const uint32_t vtests = static_cast<uint32_t>(__shfl(static_cast<int>(tests), static_cast<int>(sl), kWave));
            const uint32_t vlocal = V.src - net_first;
            const double r = nav ? draw_to_uniform(pick_draw(seed, vlocal, step, j, 1U), vnext1) : 0.0;
            uint32_t nfirst = root_first, ncur = root_nc, nslot = 0;
            // the level left for the children of an expanded child without a table slot of its own
            uint32_t rfirst = 0, rcur = 0;
            bool below = false;
            double ncum = 0.0;
            uint32_t picked = kNone;
            // level-1 pick: top-down through the table
            for (uint32_t guard = 0; guard < 80U * (kMaxLevels + 2); guard++) {
                if (__builtin_amdgcn_ballot_w64(nav) == 0) {
                    break;
                }
                if (nav) {
                    if (ncur == 0) {
                        if (below) {
                            // the children summed without reaching the draw: back to the level above
                            nfirst = rfirst;
                            ncur = rcur;
                            below = false;
                        } else {
                            nav = false;   // unreachable for a consistent table
                        }
                    } else {
                        ncur--;
                        const uint32_t y = nfirst + ncur;
                        const uint2 mt = node_meta(G, y);
                        const uint4 rq = node_rec(G, y);
                        // the child's table entry, read with its records (it is used if the lane
                        // expands the child), so that an expanded child costs one more memory round
                        // trip (its stored sum) rather than two
                        const uint32_t cs = below ? kNone : tchild[static_cast<uint64_t>(nslot) * 8 + ncur];
                        __asm__ volatile("" ::"v"(mt.x), "v"(mt.y), "v"(rq.x), "v"(rq.y), "v"(rq.z), "v"(rq.w), "v"(cs));
                        exec++;
                        const NodeEval e = eval_walk<kMu0>(s, G, V, y, mt, rq, true);
                        if (e.st == kAccept) {
                            const double cn = ncum + e.p;
                            if (reaches(cn, r)) {
                                picked = y;
                                nav = false;
                            }
                            ncum = cn;
                        } else if (e.st == kExpand) {
                            if (cs == kNoGrandSlots) {
                                // no child of this child was expanded: its children, in list order
                                rfirst = nfirst;
                                rcur = ncur;
                                below = true;
                                nfirst = mt.y & kLinkMask;
                                ncur = mt.y >> kLinkBits;
                            } else if (cs >= a.cap) {
                                nav = false;   // unreachable for a consistent table
                            } else {
                                const double ce = tcum[static_cast<uint64_t>(cs) * kWave + sl];
                                if (reaches(ce, r)) {
                                    nslot = cs;
                                    nfirst = mt.y & kLinkMask;
                                    ncur = mt.y >> kLinkBits;
                                } else {
                                    ncum = ce;
                                }
                            }
                        }
                    }
                }
            }
            // descent below the picked node
            const bool drew = item < ndraws;
            uint32_t node = picked, level = 2, dtests = vtests;
            int state = kDescFailed;
            if (drew && picked != kNone) {
                state = descend<kMu0>(s, G, V, seed, vlocal, step, j, node, level, dtests, exec);
            }
            const uint64_t sm = __builtin_amdgcn_ballot_w64(drew && state == kDescSlow);
            unsigned long long at0 = 0;
            if (ln == 0 && sm != 0) {
                at0 = atomicAdd(a.n_slow, static_cast<unsigned long long>(__popcll(sm)));
            }
            at0 = __shfl(at0, 0, kWave);
            if (drew) {
                const uint64_t slot = s.slot_off[V.src] + j;
                s.slot_tests[slot] = dtests;
                if (state == kDescSlow) {
                    a.slow[at0 + rank_below(sm)] = make_uint4(static_cast<uint32_t>(slot), V.src, node, level);
                } else {
                    s.slot_target[slot] = state == kDescDone ? node : kNone;
                }
            }
        }
        acc_exec += exec;
    }
    // executed tests of the wave, summed over the lanes
    unsigned long long v = acc_exec;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, kWave);
    }
    if (ln == 0 && v != 0) {
        atomicAdd(&s.counters[kCtrTestsExec], v);
        if (kCount) {
            atomicAdd(&s.counters[kCtrWalkSteps], cnt_steps);
            atomicAdd(&s.counters[kCtrWalkSlots], cnt_slots);
            atomicAdd(&s.counters[kCtrWalkBytes], cnt_bytes);
        }
    }
}

// Per order position i: active source of type t -> flag[t n + i] = 1.
__global__ void k_walk_flags(uint64_t n, const uint32_t* __restrict__ order, const uint8_t* __restrict__ enabled,
                             const uint32_t* __restrict__ vac, const uint8_t* __restrict__ signal,
                             const unsigned long long* __restrict__ counters, uint32_t* __restrict__ flag) {
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        const uint32_t src = order[i];
        const bool act = counters[kCtrOverCap] == 0 && enabled[src] != 0 && vac[src] != 0;
        const bool inh = signal[src] != 0;
        flag[i] = act && !inh ? 1U : 0U;
        flag[n + i] = act && inh ? 1U : 0U;
    }
}

__global__ void k_walk_scatter(uint64_t n, const uint32_t* __restrict__ flag, const uint64_t* __restrict__ off,
                               const uint32_t* __restrict__ order, uint32_t* __restrict__ list) {
    const uint64_t k = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (k < 2 * n && flag[k] != 0) {
        list[off[k]] = order[k < n ? k : k - n];
    }
}

// Segments (type, network): first list position, source count and group count.
__global__ void k_walk_segments(uint64_t n, uint32_t networks, const uint64_t* __restrict__ net_off,
                                const uint64_t* __restrict__ off, uint64_t* __restrict__ seg_first,
                                uint64_t* __restrict__ seg_count, uint32_t* __restrict__ seg_groups) {
    const uint64_t nseg = 2ULL * networks;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t sg = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; sg < nseg; sg += stride) {
        const uint64_t t = sg / networks, g = sg - t * networks;
        const uint64_t f = off[t * n + net_off[g]];
        const uint64_t e = off[t * n + net_off[g + 1]];
        seg_first[sg] = f;
        seg_count[sg] = e - f;
        seg_groups[sg] = static_cast<uint32_t>((e - f + kGroup - 1) / kGroup);
    }
}

}  // namespace

// Group-walk buffers of an FP32-mixed context: the per-update lists (sized by n and the networks) and
// the per-wave tables (sized for the waves one launch can use).
int ensure_walk(rl_bh_ctx* c) {
    const uint64_t n = c->n;
    const uint64_t nseg = 2ULL * std::max<uint32_t>(c->num_networks, 1U);
    RL_TRY(dalloc(c, &c->walk_flag, 2 * n));
    RL_TRY(dalloc(c, &c->walk_off, 2 * n + 1));
    RL_TRY(dalloc(c, &c->walk_list, n));
    RL_TRY(dalloc(c, &c->walk_seg_groups, nseg));
    RL_TRY(dalloc(c, &c->walk_seg_goff, nseg + 1));
    RL_TRY(dalloc(c, &c->walk_seg_first, nseg));
    RL_TRY(dalloc(c, &c->walk_seg_count, nseg));
    // tables for at most one wave per group, and at most about 8 KiB of table per neuron (at least
    // 256 MiB): a context of a small network does not take the tables of the whole GPU, so that many
    // such contexts fit on one device (its waves then take several groups each)
    const uint64_t groups_max = (n + kGroup - 1) / kGroup + nseg;
    const uint64_t per_wave = static_cast<uint64_t>(kWalkCap) * (kWave * sizeof(double) + 8 * sizeof(uint32_t));
    const uint64_t by_memory = std::max<uint64_t>(256ULL << 20, n * 8192ULL) / per_wave;
    c->walk_blocks_cap = static_cast<uint32_t>(
        std::min<uint64_t>(std::min<uint64_t>(static_cast<uint64_t>(c->cus) * kWalkWavesPerCu, by_memory),
                           std::max<uint64_t>(groups_max, 1)));
    c->walk_cap = kWalkCap;
    const size_t slots = static_cast<size_t>(c->walk_blocks_cap) * c->walk_cap;
    RL_TRY(dalloc(c, &c->walk_cum, slots * kWave));
    RL_TRY(dalloc(c, &c->walk_child, slots * 8));
    return RL_OK;
}

// Group walk over the active sources: lists, the walk, the per-lane general descents it hands over,
// then the per-source finalisation.  Sources handed to the per-lane kernel are in c->coop_overflow,
// counted in counters[kCtrOverflow].  All counts stay on the device.
int launch_search_walk(rl_bh_ctx* c, const SearchArgs& sa) {
    hipStream_t s = c->stream;
    const uint64_t n = c->n;
    const uint32_t G = std::max<uint32_t>(c->num_networks, 1U);
    k_walk_flags<<<grid_for(n), kBlock, 0, s>>>(n, c->order, c->enabled, c->vac[0], c->signal, c->counters,
                                                 c->walk_flag);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, c->walk_flag, 2 * n, c->walk_off));
    k_walk_scatter<<<grid_for(2 * n), kBlock, 0, s>>>(n, c->walk_flag, c->walk_off, c->order, c->walk_list);
    RL_HIP(c, hipGetLastError());
    k_walk_segments<<<grid_for(2ULL * G), kBlock, 0, s>>>(n, G, c->net_off, c->walk_off, c->walk_seg_first,
                                                           c->walk_seg_count, c->walk_seg_groups);
    RL_HIP(c, hipGetLastError());
    RL_TRY(scan_counts(c, c->walk_seg_groups, 2ULL * G, c->walk_seg_goff));
    WalkArgs a{};
    a.s = sa;
    a.list = c->walk_list;
    a.seg_goff = c->walk_seg_goff;
    a.seg_first = c->walk_seg_first;
    a.seg_count = c->walk_seg_count;
    a.nseg = 2 * G;
    a.networks = G;
    a.cursor = c->coop_cursor;
    a.chunk = std::max<uint32_t>(c->search_chunk, 1U);
    a.t_cum = c->walk_cum;
    a.t_child = c->walk_child;
    a.cap = c->walk_cap;
    a.overflow = c->coop_overflow;
    a.n_overflow = c->counters + kCtrOverflow;
    a.slow = c->slow_items;
    a.n_slow = c->counters + kCtrSlow;
    // both kernels are launched; the active-source count on the device decides which one searches
    const uint64_t* n_active = c->walk_off + 2 * n;
    a.n_active = n_active;
    a.run_from = walk_min_sources(c);
    const unsigned blocks = std::max<uint32_t>(1U, std::min<uint32_t>(c->coop_blocks, c->walk_blocks_cap));
    k_pack_walk<<<grid_for(c->tree.m), kBlock, 0, s>>>(c->tree.m, n_active, a.run_from, c->tree.meta[0],
                                                        c->tree.meta[1], c->tree.rec32[0], c->tree.rec32[1],
                                                        c->tree.wrec[0], c->tree.wrec[1]);
    RL_HIP(c, hipGetLastError());
    const bool timed = !c->capturing;
    if (timed) RL_HIP(c, hipEventRecord(c->ev2, s));
    if (a.run_from > 0) {
        // the wave-per-source kernel over the same lists (sources of each type in Morton order), at the
        // setting its tuning picks for updates below the threshold (20 waves per CU, 8 sources per grab)
        RL_TRY(launch_level1(c, sa, c->walk_list, n_active, a.run_from, c->cus * 20U, 8U));
    }
    // the counting instantiation is a separate kernel: the counters change the register allocation of the walk
    const bool count = (c->cfg.flags & RL_FLAG_COUNT_WALK) != 0;
    if (c->cfg.mu == 0.0) {
        if (count) {
            k_walk<true, true><<<blocks, kWave, 0, s>>>(a);
        } else {
            k_walk<true, false><<<blocks, kWave, 0, s>>>(a);
        }
    } else {
        if (count) {
            k_walk<false, true><<<blocks, kWave, 0, s>>>(a);
        } else {
            k_walk<false, false><<<blocks, kWave, 0, s>>>(a);
        }
    }
    RL_HIP(c, hipGetLastError());
    if (timed) RL_HIP(c, hipEventRecord(c->ev3, s));
    RL_TRY(launch_descend_lane(c, sa, c->slow_items, c->counters + kCtrSlow));
    return launch_finalize(c, sa);
}

}  // namespace rl
