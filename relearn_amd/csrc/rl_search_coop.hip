// Wavefront-cooperative Barnes-Hut target search.
//
// One wavefront handles one source neuron at a time.  The lanes form 8 groups of 8; in the
// traversal a group takes one expanded node and each of its lanes tests one child, so a wavefront
// tests the children of 8 nodes per step.  The traversal is breadth-first over exactly the nodes
// that BarnesHutBase::get_nodes_to_consider (source/algorithm/BarnesHutInternal/BarnesHutBase.h:
// 129-206) tests -- the children of every expanded node, each with test_acceptance_criterion
// (:72-114) -- so the acceptance-test count equals the CPU path's.
//
// Expanded nodes get a slot in a per-wave table: node, first expanded-child slot and count, and
// the node's mass = total probability of the accepted candidates below it.  A group sums the
// probabilities of its accepted children with a DPP butterfly; after the traversal the masses of
// expanded children are pulled into their parents, deepest level first.  The mass of a node is the
// span its subtree occupies in the reference's candidate list (descending Morton order = children
// visited in descending octant order), so a draw r is resolved top-down
// (ProbabilityPicker::pick_target, source/util/ProbabilityPicker.h:46-112): at each expanded node
// the children are scanned in descending octant order with the running sum of their masses and the
// first child whose running sum reaches r is entered, or returned when it is an accepted candidate.
// Up to 8 draws (one per group) are resolved at once.  The descents below a picked node
// (BarnesHutBase::find_target_neuron, :223-255) usually accept all children of the node; a group
// then does the whole level -- test, sum, draw, pick -- on its own, in parallel for all axons of the
// neuron.  Levels that need expansion (rare) run through the general traversal afterwards.
//
// Sums are formed in a fixed order, so results are deterministic and independent of launch
// geometry; they differ from the CPU's left-to-right sums only by rounding, so only draws within
// ~1e-15 of a cumulative boundary can resolve differently.
//
// Work distribution: persistent waves; active sources in Morton order are split into one
// contiguous segment per XCD so that an XCD's L2 caches one region of the tree; waves take sources
// from their XCD's segment, then from the others.  A source whose slot table would overflow is handed
// to the per-lane kernel.
#include <algorithm>

#include "rl_search_eval.h"

namespace rl {
namespace {

constexpr int kWave = 64;
constexpr uint32_t kQueueCap = 1024;   // LDS ring of queued / not yet flushed slots (power of two)
constexpr uint32_t kStageCap = 256;    // LDS ring of slot records not yet flushed (power of two)
constexpr uint32_t kFlushAt = 192;     // flush when this many processed slots are staged
constexpr int kXcds = kCursorSlots;
constexpr uint32_t kSlotNone = 0x0FFFFFFFU;

struct CoopArgs {
    SearchArgs s;
    const uint32_t* work;
    const uint64_t* nwork;        // device count of work; XCD segment x is [nwork x / 8, nwork (x + 1) / 8)
    unsigned long long* cursor;   // kXcds counters
    uint32_t chunk;               // sources taken from a segment per cursor increment
    uint32_t* t_my;               // per wave: cap slots, packed child range of the slot's node
    uint32_t* t_fexp;             // first expanded child slot | expanded-child count << 28
    double* t_mass;
    uint32_t cap;
    uint32_t* overflow;
    unsigned long long* n_overflow;
    uint4* slow;                  // (slot, source, node, level) of descents for the per-lane kernel
    unsigned long long* n_slow;
    uint64_t run_below;           // the kernel returns at once when *nwork >= run_below
};

// Slot records are built in LDS during the traversal and copied to the wave's global slot table in
// batches: a global store inside the step loop would make every step wait for the previous step's
// stores (loads and stores share one wait counter).  The last entry of each ring takes the writes
// of lanes with nothing to write.
struct CoopShared {
    uint32_t queue[kQueueCap + 1];     // packed child range (meta.y) of slot i at i % kQueueCap
    uint32_t stage_fx[kStageCap + 1];  // first expanded child slot | count << 28 of slot i at i % kStageCap
    double stage_ms[kStageCap + 1];    // probability mass of slot i at i % kStageCap
    uint32_t lvl_start[kMaxLevels + 4];
};

__device__ __forceinline__ uint32_t lane() {
    return __lane_id();
}

// Set bits of `mask` below the calling lane (v_mbcnt_lo/hi: two instructions instead of a masked
// 64-bit population count).
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32U),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0U));
}

__device__ __forceinline__ double uniform_f64(double v) {
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Lanes whose DPP source lies outside their row read 0 (bound_ctrl).
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// Sum over the 8 lanes of a group, valid in the group's first lane: ((p0+p1)+(p2+p3))+((p4+p5)+(p6+p7)).
__device__ __forceinline__ double group_sum8(double v) {
    v = v + dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v = v + dpp_f64<0x104>(v);  // row_shl:4
    return v;
}

// Running sum in descending child order within a group: lane j gets sum of lanes j..7 of its group.
__device__ __forceinline__ double group_suffix8(double v) {
    const uint32_t j = lane() & 7U;
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
        const double t = __shfl_down(v, d, kWave);
        if (j + d < 8) {
            v += t;
        }
    }
    return v;
}

// Element i of a wave-uniform array whose byte offsets fit 32 bits (slot tables; node arrays of
// fewer than 2^28 nodes): the address is a scalar base plus a 32-bit lane offset.
template <typename T>
__device__ __forceinline__ T& at(T* base, uint32_t i) {
    const uint32_t off = i * static_cast<uint32_t>(sizeof(T));
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}

template <typename T>
__device__ __forceinline__ const T& at(const T* base, uint32_t i) {
    const uint32_t off = i * static_cast<uint32_t>(sizeof(T));
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

struct NodeData {
    uint2 m;
    uint4 r32;
    double4 r64;
};

// Loads the count/link word and the position record of `node` (both issued together).  The loads
// are issued in every lane (node 0 stands in for inactive lanes) so that no branch hides them from
// the compiler's memory-counter bookkeeping: a load in a skippable branch makes the next wait drain
// every outstanding load, prefetches included.
template <bool kMixed, bool kMu0>
__device__ __forceinline__ NodeData load_node(const Lane& L, uint32_t node, bool active) {
    const uint32_t x = active ? node : 0U;
    NodeData d;
    d.m = at(L.meta, x);
    if (kMixed) {
        d.r32 = at(L.rec32, x);
        d.r64 = make_double4(0.0, 0.0, 0.0, 0.0);
    } else {
        d.r32 = make_uint4(0U, 0U, 0U, 0U);
        d.r64 = L.rec64[x];
    }
    return d;
}

// Acceptance test and probability of a loaded node (test_acceptance_criterion, BarnesHutBase.h:72-114;
// Kernel::calculate_attractiveness_to_connect, Kernel.h:61-75) for the lanes with `act`; other lanes
// return kDiscard and p = 0.  The FP32 path is branch-free; exact coincidences and decisions inside
// the guard band are redone in FP64 under one wave-uniform branch.
template <bool kMixed, bool kMu0>
__device__ __forceinline__ int eval_node(const SearchArgs& a, const Lane& L, const NodeData& d, uint32_t node, bool act,
                                         double* p_out) {
    // The source's own leaf lies at distance 0 and is discarded by the test, so the kernel's
    // self-exclusion (Kernel.h:61-75) never applies here.
    const bool valid = act && d.m.x != 0;
    const bool leaf = (d.m.y >> kLinkBits) == 0;
    if (kMixed) {
        const int32_t ix = static_cast<int32_t>(d.r32.x) - L.qx;
        const int32_t iy = static_cast<int32_t>(d.r32.y) - L.qy;
        const int32_t iz = static_cast<int32_t>(d.r32.z) - L.qz;
        const float dx = static_cast<float>(ix);
        const float dy = static_cast<float>(iy);
        const float dz = static_cast<float>(iz);
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        const float s2 = __uint_as_float(d.r32.w);
        const float g = __builtin_fmaf(a.theta2f, d2, -s2);
        const bool coinc = (ix | iy | iz) == 0;
        const bool near = !leaf && fabsf(g) <= __builtin_fmaf(a.guard, s2, a.guard_abs);
        int st = (leaf || g > 0.0f) ? kAccept : kExpand;
        float x;
        if (kMu0) {
            x = d2 * L.k2q;
        } else {
            const float dd = sqrtf(d2) * L.inv_q - a.muf;
            x = dd * dd * a.k2f;
        }
        const float w = __builtin_amdgcn_exp2f(x) * static_cast<float>(d.m.x);
        double p = static_cast<double>(w);
        st = valid ? st : kDiscard;
        p = st == kAccept ? p : 0.0;
        const bool need = valid && (coinc || near);
        if (__builtin_amdgcn_ballot_w64(need) != 0) {
            if (need) {
                const Eval e = eval_exact(a, L, node, 0, leaf);
                st = e.status;
                p = st == kAccept ? probability<false>(a, L, e, d.m.x, kGauss) : 0.0;
            }
        }
        *p_out = p;
        return st;
    } else {
        double p = 0.0;
        int st = kDiscard;
        if (valid) {
            const Eval e = eval_exact_rec(a, L, d.r64, leaf);
            st = e.status;
            if (st == kAccept) {
                p = probability<false>(a, L, e, d.m.x, kGauss);
            }
        }
        *p_out = p;
        return st;
    }
}

struct Bfs {
    double total;
    uint32_t count;
    uint32_t tests;
    bool overflow;
};


// One BFS step's input: up to 8 queue entries [kb, end) (one per group) and this lane's child of
// its group's node (loaded).
struct StepIn {
    uint32_t kb;     // first queue entry (= slot) of the step
    uint32_t end;    // one past the last entry the step covers
    uint32_t k;      // this group's entry
    uint32_t child;
    bool gact;       // the group has an entry
    bool act;        // the lane has a child
    NodeData nd;
};

// Reads the entries [kb, min(kb + 8, tail)) of the queue and issues the loads of their children.
template <bool kMixed, bool kMu0>
__device__ __forceinline__ StepIn fetch_step(const Lane& L, const CoopShared& sh, uint32_t kb, uint32_t tail) {
    const uint32_t ln = lane();
    const uint32_t q = ln >> 3U, j = ln & 7U;
    StepIn x;
    // kb and end are wave-uniform; keeping them in scalar registers keeps the step's control flow scalar
    x.kb = __builtin_amdgcn_readfirstlane(kb);
    x.end = __builtin_amdgcn_readfirstlane(min(kb + 8U, tail));
    x.k = kb + q;
    x.gact = x.k < x.end;
    const uint32_t packed = sh.queue[x.gact ? (x.k & (kQueueCap - 1U)) : kQueueCap];
    x.child = (packed & kLinkMask) + j;
    x.act = x.gact && j < (packed >> kLinkBits);
    x.nd = load_node<kMixed, kMu0>(L, x.child, x.act);
    return x;
}

struct SlotTable {
    uint32_t* my;
    uint32_t* fx;
    double* ms;
    uint32_t dump;   // last slot: takes the stores of lanes with nothing to store
};

struct BfsState {
    uint32_t nslots;   // queue tail = slots allocated
    uint32_t lvl_end;  // one past the last slot of the level being processed
    uint32_t lev;      // that level
    uint32_t flushed;  // slots below this are in the global table
};

// Copies the records of slots [from, to) from the LDS rings to the global slot table.
__device__ __forceinline__ void flush_slots(const SlotTable& t, const CoopShared& sh, uint32_t from, uint32_t to) {
    for (uint32_t sl = from + lane(); sl < to; sl += kWave) {
        at(t.my, sl) = sh.queue[sl & (kQueueCap - 1U)];
        at(t.fx, sl) = sh.stage_fx[sl & (kStageCap - 1U)];
        at(t.ms, sl) = sh.stage_ms[sl & (kStageCap - 1U)];
    }
}

// Tests the children of one step's entries, sums the accepted probabilities per group, appends the
// expanded children to the queue and records the entries' masses.  Entries of two consecutive levels
// can share a step; the level boundary is tracked for the upward mass pass.  Returns false when the
// slot table, the queue or the level table would overflow.
template <bool kMixed, bool kMu0>
__device__ __forceinline__ bool bfs_step(const CoopArgs& a, const Lane& L, CoopShared& sh, const SlotTable& t,
                                         const StepIn& x, BfsState& b, Bfs& r) {
    const uint32_t ln = lane();
    const uint32_t q = ln >> 3U, j = ln & 7U;
    double p;
    const int st = eval_node<kMixed, kMu0>(a.s, L, x.nd, x.child, x.act, &p);
    r.tests += static_cast<uint32_t>(__popcll(__builtin_amdgcn_ballot_w64(x.act)));
    r.count += static_cast<uint32_t>(__popcll(__builtin_amdgcn_ballot_w64(st == kAccept)));
    const double gsum = group_sum8(p);
    const bool ex = st == kExpand;
    const uint64_t em = __builtin_amdgcn_ballot_w64(ex);
    const uint32_t ne = static_cast<uint32_t>(__popcll(em));
    if (b.nslots + ne > t.dump || b.nslots + ne - b.flushed > kQueueCap) {
        return false;
    }
    // lane 8q (the one that writes group q's record) has rank_below(em) = the group's first slot
    const uint32_t slot = b.nslots + rank_below(em);
    sh.queue[ex ? (slot & (kQueueCap - 1U)) : kQueueCap] = x.nd.m.y;
    const bool gw = x.gact && j == 0;
    const uint32_t gem = static_cast<uint32_t>((em >> (8U * q)) & 0xFFULL);
    const uint32_t gi = gw ? (x.k & (kStageCap - 1U)) : kStageCap;
    sh.stage_ms[gi] = gsum;
    sh.stage_fx[gi] = (gem != 0 ? slot : kSlotNone) | (static_cast<uint32_t>(__popc(gem)) << 28U);
    // Levels that end inside this step: the level after each is the children of the groups before its
    // end.  A step can close two levels when the second was queued entirely before the step (the
    // last level of the traversal); an empty next level ends the traversal.  All of this is
    // wave-uniform scalar work.
    const uint32_t kb = __builtin_amdgcn_readfirstlane(x.kb);
    const uint32_t end = __builtin_amdgcn_readfirstlane(x.end);
    const uint32_t ns = __builtin_amdgcn_readfirstlane(b.nslots);
    uint32_t le = __builtin_amdgcn_readfirstlane(b.lvl_end);
    uint32_t lev = __builtin_amdgcn_readfirstlane(b.lev);
    for (int pass = 0; pass < 2; pass++) {
        if (!(kb < le && le <= end)) {
            break;
        }
        const uint32_t gb = le - kb;
        const uint64_t before = gb >= 8U ? ~0ULL : ((1ULL << (8U * gb)) - 1ULL);
        lev++;
        if (lev + 1 >= kMaxLevels + 4) {
            return false;
        }
        if (ln == 0) {
            sh.lvl_start[lev] = le;
        }
        const uint32_t next_end = ns + static_cast<uint32_t>(__popcll(em & before));
        if (next_end == le) {
            break;
        }
        le = next_end;
    }
    b.lvl_end = le;
    b.lev = lev;
    b.nslots = ns + ne;
    if (x.end - b.flushed >= kFlushAt) {
        flush_slots(t, sh, b.flushed, x.end);
        b.flushed = __builtin_amdgcn_readfirstlane(x.end);
    }
    return true;
}

// Breadth-first traversal below `root` (inner, with vacant dendrites) filling the slot table at `base`.
// The queue is consumed 8 entries per step across level boundaries; the next step's child records are
// loaded before the current step is evaluated (two alternating step buffers, so no registers are
// copied).
template <bool kMixed, bool kMu0>
__device__ Bfs bfs(const CoopArgs& a, const Lane& L, uint32_t root_meta_y, CoopShared& sh, uint64_t base) {
    const uint32_t ln = lane();
    Bfs r{ 0.0, 0, 0, false };
    SlotTable t;
    t.my = a.t_my + base;
    t.fx = a.t_fexp + base;
    t.ms = a.t_mass + base;
    t.dump = a.cap - 1;
    if (ln == 0) {
        sh.queue[0] = root_meta_y;
        sh.lvl_start[0] = 0;
    }
    __syncthreads();
    BfsState b{ 1U, 1U, 0U, 0U };
    StepIn xa = fetch_step<kMixed, kMu0>(L, sh, 0U, 1U);
    // Wait for the first step's data here, so that the loop header carries no pending load and the
    // stores of the previous step are never waited for.
    if (kMixed) {
        __asm__ volatile("" ::"v"(xa.nd.m.x), "v"(xa.nd.m.y), "v"(xa.nd.r32.x), "v"(xa.nd.r32.y), "v"(xa.nd.r32.z),
                         "v"(xa.nd.r32.w));
    } else {
        __asm__ volatile("" ::"v"(xa.nd.m.x), "v"(xa.nd.m.y), "v"(xa.nd.r64.x), "v"(xa.nd.r64.y), "v"(xa.nd.r64.z),
                         "v"(xa.nd.r64.w));
    }
    while (true) {
        StepIn xb = fetch_step<kMixed, kMu0>(L, sh, xa.end, b.nslots);
        __builtin_amdgcn_sched_barrier(0);   // keep the prefetch ahead of this step's work
        if (!bfs_step<kMixed, kMu0>(a, L, sh, t, xa, b, r)) {
            r.overflow = true;
            return r;
        }
        if (xb.kb >= b.nslots) {
            break;
        }
        xa = fetch_step<kMixed, kMu0>(L, sh, xb.end, b.nslots);
        __builtin_amdgcn_sched_barrier(0);
        if (!bfs_step<kMixed, kMu0>(a, L, sh, t, xb, b, r)) {
            r.overflow = true;
            return r;
        }
        if (xa.kb >= b.nslots) {
            break;
        }
    }
    const uint32_t lev = b.lev;
    flush_slots(t, sh, b.flushed, b.nslots);
    __syncthreads();
    // pull the masses of expanded children into their parents, deepest parents first
    for (int lv = static_cast<int>(lev) - 2; lv >= 0; lv--) {
        const uint32_t sb = sh.lvl_start[lv], se = sh.lvl_start[lv + 1];
        for (uint32_t i0 = sb; i0 < se; i0 += kWave) {
            const uint32_t sl = i0 + ln;
            if (sl < se) {
                const uint32_t fx = at(t.fx, sl);
                const uint32_t ne = fx >> 28U;
                const uint32_t f = fx & kSlotNone;
                if (ne != 0 && f + ne <= t.dump) {
                    double m8[8];
#pragma unroll
                    for (uint32_t c = 0; c < 8; c++) {
                        m8[c] = at(t.ms, f + (c < ne ? c : 0U));
                    }
                    double sum = 0.0;
#pragma unroll
                    for (uint32_t c = 0; c < 8; c++) {
                        if (c < ne) {
                            sum += m8[c];
                        }
                    }
                    at(t.ms, sl) += sum;
                }
            }
        }
        __syncthreads();
    }
    r.total = t.ms[0];
    return r;
}

// Resolves the draws of the active groups top-down through the slot table (one draw per group).
// Returns the picked node in every lane of the group, kNone when the list had no positive mass.
template <bool kMixed, bool kMu0>
__device__ uint32_t navigate(const CoopArgs& a, const Lane& L, double rr, bool gact, uint64_t base, uint32_t* tests) {
    const SearchArgs& s = a.s;
    const uint32_t ln = lane();
    const uint32_t q = ln >> 3U, j = ln & 7U;
    const uint32_t g0 = 8U * q;
    const uint32_t* const tmy = a.t_my + base;
    const uint32_t* const tfx = a.t_fexp + base;
    const double* const tms = a.t_mass + base;
    uint32_t slot = 0;
    bool done = !gact;
    uint32_t res = kNone;
    for (uint32_t guard = 0; guard < kMaxLevels + 2; guard++) {
        if (__ballot(!done) == 0) {
            break;
        }
        // slot < cap always holds: it only advances to a checked child slot
        uint32_t fx = at(tfx, slot);
        uint32_t my = at(tmy, slot);
        if (done) {
            fx = 0;
            my = 0;
        }
        const uint32_t fc = my & kLinkMask;
        const uint32_t nc = my >> kLinkBits;
        const uint32_t f = fx & kSlotNone;
        const uint32_t nexp = fx >> 28U;
        const bool act = !done && j < nc;
        // the children and the masses of the expanded ones, loaded together
        const NodeData nd = load_node<kMixed, kMu0>(L, fc + j, act);
        const bool mact = !done && j < nexp && f != kSlotNone && f + j < a.cap;
        const double mj_raw = at(tms, mact ? f + j : 0U);
        const double mj = mact ? mj_raw : 0.0;
        double p;
        const int st = eval_node<kMixed, kMu0>(s, L, nd, fc + j, act, &p);
        *tests += static_cast<uint32_t>(__popcll(__ballot(act)));
        const uint64_t em = __ballot(act && st == kExpand);
        const uint32_t gem = static_cast<uint32_t>((em >> g0) & 0xFFULL);
        const uint32_t rk = static_cast<uint32_t>(__popc(gem & ((1U << j) - 1U)));
        const double mexp = __shfl(mj, static_cast<int>(g0 + (rk & 7U)), kWave);
        double mass = 0.0;
        if (st == kAccept) {
            mass = p;
        } else if (st == kExpand) {
            mass = mexp;
        }
        const double cum = group_suffix8(mass);
        const double cum_next_raw = __shfl_down(cum, 1, kWave);
        const double cum_next = (j + 1 < 8) ? cum_next_raw : 0.0;
        const uint32_t gcm = static_cast<uint32_t>((__ballot(act && mass > 0.0 && cum >= rr) >> g0) & 0xFFULL);
        const uint32_t gpm = static_cast<uint32_t>((__ballot(act && mass > 0.0) >> g0) & 0xFFULL);
        int jstar = -1;
        if (gcm != 0) {
            jstar = 31 - __clz(static_cast<int>(gcm));
        } else if (gpm != 0) {
            jstar = __ffs(static_cast<int>(gpm)) - 1;
        }
        const uint32_t src_lane = g0 + static_cast<uint32_t>(jstar < 0 ? 0 : jstar);
        const int stj = __shfl(st, static_cast<int>(src_lane), kWave);
        const double before = __shfl(cum_next, static_cast<int>(src_lane), kWave);
        if (!done) {
            if (jstar < 0) {
                done = true;
                res = kNone;
            } else if (stj == kAccept) {
                done = true;
                res = fc + static_cast<uint32_t>(jstar);
            } else {
                rr = rr - before;
                const uint32_t next = f + static_cast<uint32_t>(__popc(gem & ((1U << jstar) - 1U)));
                if (f == kSlotNone || next >= a.cap - 1) {
                    done = true;
                    res = kNone;
                } else {
                    slot = next;
                }
            }
        }
    }
    return res;
}

struct ListResult {
    double total;
    uint32_t count;
    uint32_t tests;
    bool overflow;   // slot table overflow or a Gaussian sum of 0: the per-lane kernel redoes the source
};

// Candidate list of the subtree of `root`; a list whose Gaussian sum underflows to 0 needs the
// reference's fallbacks (Kernel::create_probability_interval, Kernel.h:109-126), which the per-lane
// kernel implements, so the source is handed over.
template <bool kMixed, bool kMu0>
__device__ ListResult list_of(const CoopArgs& a, const Lane& L, uint32_t root, CoopShared& sh, uint64_t base) {
    ListResult lr{ 0.0, 0, 0, false };
    const uint2 mr = L.meta[root];
    if (mr.x == 0) {
        return lr;
    }
    const Bfs b = bfs<kMixed, kMu0>(a, L, mr.y, sh, base);
    lr.tests = b.tests;
    lr.total = b.total;
    lr.count = b.count;
    lr.overflow = b.overflow || (b.count > 0 && b.total == 0.0);
    return lr;
}


// Per-draw descent state, identical in the 8 lanes of the draw's group.
enum : int { kDrawActive = 0, kDrawDone = 1, kDrawFailed = 2, kDrawSlow = 3, kDrawIdle = 4 };

// Levels >= 2 of the groups' draws (BarnesHutBase::find_target_neuron, :223-255), one draw per group
// of 8 lanes, each lane testing one child of the draw's current node.  Below an accepted node the
// candidate list is normally the node's children, all accepted; the group then sums, draws and picks
// on its own.  A level that needs expansion, or whose Gaussian sum is 0 (the reference's fallbacks),
// ends in kDrawSlow with (node, level) where the general per-lane traversal takes over.
template <bool kMixed, bool kMu0>
__device__ __forceinline__ int descend_fast(const SearchArgs& s, const Lane& L, uint32_t jd, int state, uint32_t& node,
                                            uint32_t& level, uint32_t& tests, unsigned long long& exec) {
    const uint32_t ln = lane();
    const uint32_t j = ln & 7U;
    const uint32_t g0 = ln & ~7U;
    for (uint32_t guard = 0; guard < kMaxLevels + 2; guard++) {
        uint32_t my = 0;
        if (state == kDrawActive) {
            if (node >= s.m) {
                state = kDrawFailed;
            } else {
                const uint2 mn = L.meta[node];
                my = mn.y;
                if ((my >> kLinkBits) == 0) {
                    node = my & kLinkMask;  // leaf: the target neuron
                    state = kDrawDone;
                } else if (mn.x == 0) {
                    state = kDrawFailed;  // no vacant dendrites: empty list
                }
            }
        }
        if (__builtin_amdgcn_ballot_w64(state == kDrawActive) == 0) {
            break;
        }
        const bool act_g = state == kDrawActive;
        const uint32_t fc = my & kLinkMask;
        const uint32_t nc = act_g ? (my >> kLinkBits) : 0U;
        const bool act = act_g && j < nc;
        double p;
        const NodeData cd = load_node<kMixed, kMu0>(L, fc + j, act);
        const int st = eval_node<kMixed, kMu0>(s, L, cd, fc + j, act, &p);
        exec += static_cast<uint32_t>(__popcll(__builtin_amdgcn_ballot_w64(act)));
        const uint32_t gexp = static_cast<uint32_t>((__builtin_amdgcn_ballot_w64(st == kExpand) >> g0) & 0xFFULL);
        const uint32_t gacc = static_cast<uint32_t>((__builtin_amdgcn_ballot_w64(st == kAccept) >> g0) & 0xFFULL);
        const double gsum = __shfl(group_sum8(p), static_cast<int>(g0), kWave);
        if (act_g) {
            if (gexp != 0 || (gacc != 0 && gsum == 0.0)) {
                state = kDrawSlow;
            } else {
                tests += nc;
                if (gacc == 0) {
                    state = kDrawFailed;
                }
            }
        }
        const bool pick = state == kDrawActive;
        const double r = pick ? draw_to_uniform(pick_draw(L.seed, L.local, L.step, jd, level), nextafter(gsum, gsum + 0.00001))
                              : 0.0;
        const double pa = st == kAccept ? p : 0.0;
        const double cum = group_suffix8(pa);
        const uint32_t gcm = static_cast<uint32_t>((__builtin_amdgcn_ballot_w64(pa > 0.0 && cum >= r) >> g0) & 0xFFULL);
        const uint32_t gpm = static_cast<uint32_t>((__builtin_amdgcn_ballot_w64(pa > 0.0) >> g0) & 0xFFULL);
        if (pick) {
            int jstar = -1;
            if (gcm != 0) {
                jstar = 31 - __clz(static_cast<int>(gcm));
            } else if (gpm != 0) {
                jstar = __ffs(static_cast<int>(gpm)) - 1;
            }
            if (jstar < 0) {
                state = kDrawFailed;
            } else {
                node = fc + static_cast<uint32_t>(jstar);
                level++;
            }
        }
    }
    return state;
}

// One wave per source builds the candidate list of the root once (BarnesHutBase::find_target_neuron
// re-derives it for every axon, BarnesHutBase.h:223-255), resolves the source's level-1 draws against
// it 8 at a time, and continues each draw's descent in its group.  Results go to slot_target; descents
// that need the general traversal are queued for the per-lane kernel.  Both instances are held to 5
// waves per SIMD (at most 96 VGPRs): the FP64 one needs 98 otherwise; to fit, it spills one 64-bit
// value (stored once, reloaded once per source, never in the step loop), and measures 5% faster at 20
// waves per CU than at 4 waves per SIMD.
template <bool kMixed, bool kMu0>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(5, 8))) k_level1(CoopArgs a) {
    __shared__ CoopShared sh;
    const SearchArgs& s = a.s;
    const uint32_t ln = lane();
    const uint32_t q = ln >> 3U, j = ln & 7U;
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * a.cap;
    int seg = static_cast<int>(blockIdx.x % kXcds);
    int tried = 0;
    unsigned long long acc_exec = 0;
    const uint64_t nw = *a.nwork;
    if (nw >= a.run_below) {
        return;   // the group walk takes this update
    }
    const uint32_t step = *s.step;
    const uint32_t chunk = a.chunk;
    uint64_t w_next = 0, w_lim = 0;   // the wave's current chunk of its segment
    while (true) {
        if (w_next >= w_lim) {
            const uint64_t seg_end = nw * static_cast<uint64_t>(seg + 1) / kXcds;
            uint64_t w0 = 0;
            if (ln == 0) {
                w0 = nw * static_cast<uint64_t>(seg) / kXcds + atomicAdd(&a.cursor[seg], static_cast<unsigned long long>(chunk));
            }
            w0 = __shfl(w0, 0, kWave);
            if (w0 >= seg_end) {
                if (++tried >= kXcds) {
                    break;
                }
                seg = (seg + 1) % kXcds;
                continue;
            }
            w_next = w0;
            w_lim = std::min<uint64_t>(w0 + chunk, seg_end);
        }
        const uint64_t w = w_next++;
        const uint32_t src = __builtin_amdgcn_readfirstlane(a.work[w]);
        const uint32_t na = __builtin_amdgcn_readfirstlane(s.vac_axon[src]);
        const uint32_t slot0 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(s.slot_off[src]));
        // source data is wave-uniform: keep it in scalar registers
        Lane L;
        L.src = src;
        L.sig = __builtin_amdgcn_readfirstlane(s.signal[src] == 0 ? 0 : 1);
        lane_bind(s, L);
        const uint32_t net = __builtin_amdgcn_readfirstlane(s.net_of[src]);
        L.root = net;
        L.local = src - __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(s.net_off[net]));
        L.step = step;
        L.seed = (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s.net_seed[net] >> 32U)))) << 32U) |
                 static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(s.net_seed[net] & 0xFFFFFFFFULL)));
        L.k2q = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s.net_k2q[net])));
        L.inv_q = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s.net_invq[net])));
        L.px = uniform_f64(s.pos[src]);
        L.py = uniform_f64(s.pos[s.n + src]);
        L.pz = uniform_f64(s.pos[2 * s.n + src]);
        if (kMixed) {
            L.qx = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(s.qpos[src]));
            L.qy = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(s.qpos[s.n + src]));
            L.qz = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(s.qpos[2 * s.n + src]));
        } else {
            L.qx = L.qy = L.qz = 0;
        }
        const ListResult l1 = list_of<kMixed, kMu0>(a, L, L.root, sh, base);
        acc_exec += l1.tests;
        if (l1.overflow) {
            if (ln == 0) {
                const unsigned long long k = atomicAdd(a.n_overflow, 1ULL);
                a.overflow[k] = src;
                s.found[src] = kOverflowMark;
            }
        } else if (l1.count == 0) {
            if (ln == 0) {
                s.slot_tests[slot0] = l1.tests;   // slot_target[slot0] stays kNone: one failed search
            }
        } else {
            const double next1 = nextafter(l1.total, l1.total + 0.00001);
            uint32_t nav_tests = 0;
            for (uint32_t j0 = 0; j0 < na; j0 += 8) {
                const bool gact = j0 + q < na;
                const double r = gact ? draw_to_uniform(pick_draw(L.seed, L.local, L.step, j0 + q, 1U), next1) : 0.0;
                uint32_t node = navigate<kMixed, kMu0>(a, L, r, gact, base, &nav_tests);
                uint32_t level = 2, tests = l1.tests;
                int state = !gact ? kDrawIdle : (node == kNone ? kDrawFailed : kDrawActive);
                state = descend_fast<kMixed, kMu0>(s, L, j0 + q, state, node, level, tests, acc_exec);
                const bool lead = gact && j == 0;
                const uint64_t sm = __builtin_amdgcn_ballot_w64(lead && state == kDrawSlow);
                unsigned long long at0 = 0;
                if (ln == 0 && sm != 0) {
                    at0 = atomicAdd(a.n_slow, static_cast<unsigned long long>(__popcll(sm)));
                }
                at0 = __shfl(at0, 0, kWave);
                if (lead) {
                    const uint32_t slot = slot0 + j0 + q;
                    s.slot_tests[slot] = tests;
                    if (state == kDrawSlow) {
                        a.slow[at0 + rank_below(sm)] = make_uint4(slot, src, node, level);
                    } else {
                        s.slot_target[slot] = state == kDrawDone ? node : kNone;
                    }
                }
            }
            acc_exec += nav_tests;
        }
        __syncthreads();
    }
    if (ln == 0) {
        atomicAdd(&s.counters[kCtrTestsExec], acc_exec);
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, kWave);
    }
    return v;
}

// Per source, in axon order: a search that fails ends the source's searches (find_target_neurons,
// BarnesHutBase.h:283-289).  Counts the searches, the successful ones and the reference's tests.
// Neurons in index order, grid-stride, so that neighbouring lanes read neighbouring slots (the work
// list is in Morton order, whose slots are scattered).  The totals are summed per block and added
// with one atomic per block: atomics of every wave on the same two counters serialise (they took
// 6 ms at 16M).
__global__ void __launch_bounds__(256) k_finalize(SearchArgs s) {
    __shared__ unsigned long long part[2][256 / kWave];
    unsigned long long ref = 0, searched_n = 0;
    const bool over = s.counters[kCtrOverCap] != 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t src = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; src < s.n; src += stride) {
        // the active sources of k_active_flags: enabled, with vacant axons, and no over-capacity update
        if (over || s.enabled[src] == 0 || s.vac_axon[src] == 0 || s.found[src] == kOverflowMark) {
            continue;
        }
        const uint32_t na = s.vac_axon[src];
        const uint64_t slot0 = s.slot_off[src];
        uint32_t found = 0, searched = 0;
        uint32_t jj = 0;
        for (; jj < na; jj++) {
            searched++;
            ref += s.slot_tests[slot0 + jj];
            if (s.slot_target[slot0 + jj] == kNone) {
                break;
            }
            found++;
        }
        // draws after the first failure were resolved speculatively: the reference never makes them
        for (jj = jj + 1; jj < na; jj++) {
            s.slot_target[slot0 + jj] = kNone;
            s.slot_tests[slot0 + jj] = 0;
        }
        s.found[src] = found;
        s.searched[src] = searched;
        searched_n += searched;
    }
    ref = wave_sum(ref);
    searched_n = wave_sum(searched_n);
    const uint32_t wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        part[0][wv] = ref;
        part[1][wv] = searched_n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0, sn = 0;
        for (uint32_t k = 0; k < 256 / kWave; k++) {
            r += part[0][k];
            sn += part[1][k];
        }
        if (sn != 0) {
            atomicAdd(&s.counters[kCtrTestsRef], r);
            atomicAdd(&s.counters[kCtrSearches], sn);
        }
    }
}

}  // namespace

// Cooperative search over the active sources (c->coop_work, count in sa.nwork): the search kernel,
// the per-lane general descents it hands over, then the per-source finalisation.  Sources that
// overflow the slot table (or whose root list sums to 0) go to c->coop_overflow, counted in
// counters[kCtrOverflow], for the per-lane kernel.  All counts stay on the device.
int launch_search_coop(rl_bh_ctx* c, const SearchArgs& sa) {
    const bool timed = !c->capturing;
    if (timed) RL_HIP(c, hipEventRecord(c->ev2, c->stream));
    RL_TRY(launch_level1(c, sa, c->coop_work, sa.nwork, UINT64_MAX, c->coop_blocks, c->search_chunk));
    if (timed) RL_HIP(c, hipEventRecord(c->ev3, c->stream));
    RL_TRY(launch_descend_lane(c, sa, c->slow_items, c->counters + kCtrSlow));
    return launch_finalize(c, sa);
}

// The wave-per-source kernel over work[0, *nwork) on `blocks` persistent waves taking `chunk` sources
// per cursor increment, skipped on the device when *nwork >= run_below.
int launch_level1(rl_bh_ctx* c, const SearchArgs& sa, const uint32_t* work, const uint64_t* nwork,
                  uint64_t run_below, uint32_t blocks, uint32_t chunk) {
    hipStream_t s = c->stream;
    CoopArgs a{};
    a.s = sa;
    a.work = work;
    a.nwork = nwork;
    a.run_below = run_below;
    a.cursor = c->coop_cursor;
    a.chunk = std::max<uint32_t>(chunk, 1U);
    a.t_my = c->coop_node;
    a.t_fexp = c->coop_fexp;
    a.t_mass = c->coop_mass;
    a.cap = c->coop_cap;
    a.overflow = c->coop_overflow;
    a.n_overflow = c->counters + kCtrOverflow;
    a.slow = c->slow_items;
    a.n_slow = c->counters + kCtrSlow;
    const bool mixed = c->cfg.precision == RL_PREC_FP32_MIXED;
    const bool mu0 = c->cfg.mu == 0.0;
    if (mixed && mu0) {
        k_level1<true, true><<<blocks, kWave, 0, s>>>(a);
    } else if (mixed) {
        k_level1<true, false><<<blocks, kWave, 0, s>>>(a);
    } else {
        k_level1<false, false><<<blocks, kWave, 0, s>>>(a);
    }
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

int launch_finalize(rl_bh_ctx* c, const SearchArgs& sa) {
    const uint64_t n = std::max<uint64_t>(c->n, 1);
    k_finalize<<<static_cast<unsigned>(std::min<uint64_t>((n + 255) / 256, kStreamBlocks)), 256, 0, c->stream>>>(sa);
    RL_HIP(c, hipGetLastError());
    return RL_OK;
}

}  // namespace rl
