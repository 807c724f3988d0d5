// Internal declarations shared by the HIP translation units of librelearn_bh.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/relearn_bh.h"

namespace rl {

constexpr uint32_t kMaxLevels = 42;          // octree levels below the root resolved by the 2x63-bit keys
constexpr uint32_t kLinkBits = 28;           // node / neuron index bits in the packed link word
constexpr uint32_t kLinkMask = (1U << kLinkBits) - 1U;
constexpr uint32_t kNoTarget = 0xFFFFFFFFU;
constexpr int kDrawsPerPass = 8;             // draws resolved per pick traversal
constexpr unsigned kStreamBlocks = 2048;     // grid of the grid-stride kernels sized on the device
constexpr int kCursorSlots = 8;              // work cursors of the cooperative search, one per XCD
constexpr uint32_t kMaxWavesPerCu = 24;      // slot tables are allocated for this many search waves per CU

// Device counters of one update (rl_bh_ctx::counters), zeroed at the start of every search and read
// back with one copy.
enum : int {
    kCtrTestsExec = 0,   // node tests executed
    kCtrTestsRef = 1,    // node tests the reference's traversal makes
    kCtrSearches = 2,
    kCtrCreated = 3,     // synapses created by the commit
    kCtrRequests = 4,
    kCtrSlots = 5,       // vacant axons of the update
    kCtrOverCap = 6,     // 1: more vacant axons than slots_cap, nothing was searched
    kCtrSlow = 7,        // descents queued for the per-lane traversal
    kCtrOverflow = 8,    // sources whose root list overflowed the cooperative slot table
    kCtrRedo = 9,        // sources of the per-lane path that need the fallback pass
    kCtrActive = 10,     // sources with vacant axons
    kCtrSticky = 11,     // over-capacity events (slots, network graph) since the host last read the counters
    kCtrNetOver = 12,    // 1: the created synapses did not fit the network graph, none was appended
    kCtrNetNew = 13,     // synapses appended to the network graph by the last commit
    kCtrDeleted = 14,    // synapse units deleted by the last element commit
    kCtrDelShort = 15,   // deletions that found fewer synapses than the element commit asked for
    kCtrWalkSteps = 16,  // group walk: node steps (one node for the wavefront's sources)
    kCtrWalkSlots = 17,  // group walk: expanded nodes given a table slot
    kCtrWalkBytes = 18,  // group walk: bytes of running sums and child entries stored in the tables
    kCounters = 19
};

struct Status {
    int code = RL_OK;
    std::string msg;
};

// Device octree in breadth-first order (level-major, Morton order inside a level).  Children of a
// node are contiguous.  Per dendrite type s (0 excitatory, 1 inhibitory) the search reads
//   rec64[s][i] = (centroid x, y, z, max edge)          FP64 path, 32 B
//   rec32[s][i] = (centroid x, y, z as 31-bit fixed point, squared max edge in squared fixed-point
//                  units as float bits, -infinity for a leaf)   FP32 path, 16 B
//   meta[s][i]  = (vacant count, link | nchild << 28)  8 B
// link is the first child for inner nodes and the neuron id for leaves.  The group walk's records
// (wrec) hold kNodePad more nodes than the tree: when the walk copies a node's children again (after a
// deeper level took their LDS block) it copies the 8 records from the first child on, whatever the
// node's child count, and the last node's first child may be among the last nodes.
constexpr uint64_t kNodePad = 8;
struct Tree {
    uint64_t m = 0;
    uint32_t depth = 0;
    std::vector<uint64_t> level_start;  // host copy, depth + 2 entries
    uint8_t* level = nullptr;
    uint8_t* octant = nullptr;
    double* box = nullptr;              // 6 per node: lo xyz, hi xyz
    double4* rec64[2] = { nullptr, nullptr };
    uint4* rec32[2] = { nullptr, nullptr };
    // FP32 path, group walk: per node rec32 and meta side by side (2 x 16 B, the meta pair in the low half
    // of the second), packed from them for every update the walk searches, so that one load brings
    // both; kNodePad spare nodes
    uint4* wrec[2] = { nullptr, nullptr };
    uint2* meta[2] = { nullptr, nullptr };
    uint32_t* node_net = nullptr;       // network of each node (the roots are nodes 0..networks-1)
};

// In-edges of every neuron for the synaptic input (the local in-edges of NetworkGraph), CSR by
// target: edges [off[t], off[t+1]) hold (source, weight).  Every commit appends the synapses it
// creates.  The append writes the merged graph to the second buffer and copies it back, so the
// arrays a captured period refers to never change.
struct NetGraph {
    uint64_t* off = nullptr;       // n + 1
    uint32_t* src = nullptr;       // cap
    int32_t* w = nullptr;          // cap
    uint64_t* off2 = nullptr;      // n + 1: merge buffers
    uint32_t* src2 = nullptr;
    int32_t* w2 = nullptr;
    uint32_t* ncnt = nullptr;      // n: synapses appended per target, then merged entries per target
    uint64_t* npre = nullptr;      // n + 1: prefix sums of ncnt
    uint32_t* ocnt = nullptr;      // n: out-entries per source of a deletion phase
    uint64_t* ooff = nullptr;      // n + 1
    uint2* olist = nullptr;        // cap: (entry, target) of each out-entry, grouped by source
    // Synaptic input by push: the non-zero pairs grouped by source, and per step parity the summed
    // weights that the neurons which fired pushed to their targets for the next step.
    uint64_t* out_off = nullptr;   // n + 1
    uint32_t* out_cnt = nullptr;   // n: zero between builds
    uint32_t* out_tgt = nullptr;   // cap
    int32_t* out_w = nullptr;      // cap
    int32_t* acc[2] = { nullptr, nullptr };  // n each
    uint64_t cap = 0;
    // host: an upper bound of the graph's entries, exact after a call that synchronised with the device
    // and raised by each replay without stats by the most synapses a period can append
    uint64_t entries_bound = 0;
    bool ready = false;
    // acc[parity] holds the pushes of fired_buf[parity] over the current pairs, acc[parity ^ 1] is
    // zero, and the out-edge lists match the graph
    bool syn_valid = false;
};

struct Neurons {
    double* x = nullptr;
    double* u = nullptr;
    double* ca = nullptr;
    double* target_ca = nullptr;
    bool target_uniform = false;  // every neuron has target_value: the step reads no per-neuron target
    double target_value = 0.0;
    double* input = nullptr;
    uint8_t* fired_buf[2] = { nullptr, nullptr };  // fired status of the last two steps
    uint8_t* fired = nullptr;     // fired_buf[parity]: the last step's
    int parity = 0;
    double* grown = nullptr;      // 3n
    double* delta = nullptr;      // 3n
    uint32_t* connected = nullptr;  // 3n
    uint32_t* deleted = nullptr;  // 3n
    double* net_nu = nullptr;     // growth rates per network (3 each), null: p.nu everywhere
    double* min_ca = nullptr;     // 3n per-neuron minimum calcium (type-major), null: p.eta
    double* retract = nullptr;    // 3n per-neuron vacant retract ratio, null: p.retract_ratio
    rl_step_params p{};
    bool ready = false;
};

}  // namespace rl

struct rl_bh_ctx {
    rl_bh_config cfg{};
    rl::Status st;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    double last_search_ms = 0.0;
    double last_kernel_ms = 0.0;    // the dominant search kernel alone

    uint64_t n = 0;
    double* pos = nullptr;          // SoA x[n], y[n], z[n]
    uint32_t* qpos = nullptr;       // SoA fixed-point positions (FP32 path)
    uint8_t* signal = nullptr;
    uint8_t* enabled = nullptr;
    uint32_t* vac[3] = { nullptr, nullptr, nullptr };  // axon, exc dendrite, inh dendrite
    uint32_t* conn[3] = { nullptr, nullptr, nullptr };
    uint32_t* leaf_of = nullptr;    // neuron -> leaf node
    uint32_t* order = nullptr;      // neurons in Morton order (search order)
    rl::Tree tree;
    bool prepared = false;

    // Networks of the context: one for rl_bh_init, an ensemble of independent networks for
    // rl_bh_init_networks.  Network g owns neurons [net_off[g], net_off[g+1]), the box
    // net_box[6g..6g+5] (lo xyz, hi xyz), the octree rooted at node g, its own fixed-point scale and
    // its own draw stream.
    uint32_t num_networks = 0;
    std::vector<uint64_t> h_net_off;
    std::vector<double> h_net_box;
    std::vector<double> h_net_scale;
    uint64_t* net_off = nullptr;    // device copies
    double* net_box = nullptr;
    double* net_scale = nullptr;    // fixed-point units per micrometre (2^31 - 1 across the largest extent)
    uint64_t* net_seed = nullptr;
    float* net_k2q = nullptr;       // -log2(e) / sigma^2 in squared fixed-point units
    float* net_invq = nullptr;      // micrometres per fixed-point unit
    uint32_t* net_of = nullptr;     // n

    // search results (slot layout: vacant axons per neuron at the exclusive prefix sum).  One
    // capacity, slots_cap, sizes every per-slot and per-request buffer: a neuron's requests are a
    // prefix of its slots, so there are never more requests than slots.
    uint64_t slots = 0;
    uint64_t slots_cap = 0;
    uint64_t* slot_off = nullptr;   // n + 1
    uint32_t* slot_target = nullptr;
    uint32_t* slot_tests = nullptr;
    uint32_t* found = nullptr;      // n
    uint32_t* searched = nullptr;   // n
    // dense requests in (source, axon ordinal) order
    uint64_t nreq = 0;
    uint32_t* req_target = nullptr;
    uint32_t* req_source = nullptr;
    uint8_t* req_signal = nullptr;
    uint32_t* req_j = nullptr;
    uint8_t* req_ok = nullptr;
    bool searched_valid = false;
    bool committed = false;          // the requests of the last search were committed (a second commit
                                     // would apply them again)
    uint64_t last_searches = 0;
    uint64_t last_active = 0;        // active sources of the last search
    uint64_t walk_min = 0;           // FP32-mixed: active sources from which an update takes the group walk
    uint64_t last_tests = 0;
    uint64_t last_tests_ref = 0;
    uint64_t last_lane_sources = 0;  // sources handled by the per-lane kernel in the last search
    uint64_t last_slow_descents = 0; // descents handed to the per-lane general traversal
    uint64_t last_walk_steps = 0;    // group walk: node steps and table slots of the last search
    uint64_t last_walk_slots = 0;
    uint64_t last_walk_bytes = 0;

    // persistent scratch, sized at init or with slots_cap (never inside an update)
    uint64_t* wide = nullptr;        // n + 1: offsets of the active sources
    uint64_t* req_off = nullptr;     // n + 1
    unsigned long long* counters = nullptr;  // rl::kCounters device counters of the current update
    uint32_t* d_steps = nullptr;     // [0] first neuron step of a captured period, [1] plasticity update
    uint64_t* scan_part = nullptr;   // tile sums of scan_counts
    uint64_t scan_part_cap = 0;
    // commit: requests grouped by (target, dendrite type) with a counting sort
    uint32_t* c_gcnt = nullptr;      // 2n
    uint64_t* c_goff = nullptr;      // 2n + 1
    uint64_t* c_prio = nullptr;      // slots_cap: processing priority per request
    uint64_t* c_order = nullptr;     // slots_cap: caller's processing order
    uint64_t* c_sprio = nullptr;     // slots_cap: priorities in group order
    uint32_t* c_sidx = nullptr;      // slots_cap: request indices in group order
    // cooperative search
    uint32_t* coop_flag = nullptr;   // n
    uint32_t* coop_work = nullptr;   // n
    unsigned long long* coop_cursor = nullptr;  // one work cursor per XCD
    uint32_t* coop_overflow = nullptr;  // n
    uint32_t* redo = nullptr;           // n
    uint32_t* coop_node = nullptr;   // blocks * cap
    uint32_t* coop_fexp = nullptr;
    double* coop_mass = nullptr;
    uint32_t coop_cap = 0;
    uint32_t coop_blocks = 0;        // persistent search waves: cus * waves_per_cu
    uint32_t cus = 0;
    uint32_t waves_per_cu = 16;
    uint32_t search_chunk = 1;       // sources per work-queue grab of a search wave
    // group walk (the FP32-mixed search): per order position and type the active flags, their offsets,
    // the grouped source list, the (type, network) segments, and per wave the walk's tables
    uint32_t* walk_flag = nullptr;        // 2n
    uint64_t* walk_off = nullptr;         // 2n + 1
    uint32_t* walk_list = nullptr;        // n
    uint32_t* walk_seg_groups = nullptr;  // 2 networks
    uint64_t* walk_seg_goff = nullptr;    // 2 networks + 1
    uint64_t* walk_seg_first = nullptr;   // 2 networks
    uint64_t* walk_seg_count = nullptr;   // 2 networks
    double* walk_cum = nullptr;           // walk_blocks_cap x walk_cap x 64
    uint32_t* walk_child = nullptr;       // walk_blocks_cap x walk_cap x 8
    uint32_t walk_cap = 0;
    uint32_t walk_blocks_cap = 0;
    // (slot, source, node, level) of descents that need the general per-lane traversal (slots_cap)
    uint4* slow_items = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    // captured plasticity period (rl_bh_graph_create)
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_steps = 0;
    uint64_t graph_max_vacant = 0;
    bool graph_appends = false;   // the captured period appends its synapses to the network graph
    bool capturing = false;
    rl::Neurons neurons;
    rl::NetGraph net;
};

namespace rl {

int fail(rl_bh_ctx* c, int code, const std::string& msg);
// message of the last failing context-free call of this thread (rl_bh_last_error(NULL)); a fixed
// buffer, so setting it allocates nothing
void set_global_error(const char* msg);
int hip_fail(rl_bh_ctx* c, hipError_t e, const char* what);
int ensure_tmp(rl_bh_ctx* c, size_t bytes);

template <typename T>
int dalloc(rl_bh_ctx* c, T** p, size_t count) {
    if (*p != nullptr) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) {
        count = 1;
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, RL_EOOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return RL_OK;
}

template <typename T>
void dfree(T** p) {
    if (*p != nullptr) {
        (void)hipFree(*p);
        *p = nullptr;
    }
}

// tree (rl_tree.hip)
int build_tree(rl_bh_ctx* c, const double* h_xyz);
int prepare_tree(rl_bh_ctx* c);
void free_tree(rl_bh_ctx* c);
int export_tree(rl_bh_ctx* c, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant, double* box_lo,
                double* box_hi, double* side, uint32_t* cnt_exc, uint32_t* cnt_inh, uint8_t* has_exc, uint8_t* has_inh,
                double* pos_exc, double* pos_inh);

// captured plasticity periods (rl_api.hip)
void destroy_graph(rl_bh_ctx* c);

// scans and fills (rl_scan.hip)
uint64_t scan_scratch_entries(uint64_t n);
int scan_counts(rl_bh_ctx* c, const uint32_t* in, uint64_t n, uint64_t* out);  // out[0..n], out[n] = total
// the same on a given stream with scan_scratch_entries(n) tile sums at part (16-byte aligned operands)
hipError_t scan_counts_on(hipStream_t s, const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* part);
hipError_t fill_u32_on(hipStream_t s, uint32_t* p, uint32_t value, uint64_t n, const uint64_t* n_dev, uint64_t cap);
int fill_u32(rl_bh_ctx* c, uint32_t* p, uint32_t value, uint64_t n, const uint64_t* n_dev, uint64_t cap);
int set_steps(rl_bh_ctx* c, uint32_t neuron_step, uint32_t update);

// search (rl_search.hip, rl_search_coop.hip)
int ensure_slots(rl_bh_ctx* c, uint64_t want);
int search_capacity(rl_bh_ctx* c);
int enqueue_search(rl_bh_ctx* c);
int find_targets(rl_bh_ctx* c, uint32_t step);
int read_counters(rl_bh_ctx* c, unsigned long long* h);
int candidate_lists(rl_bh_ctx* c, uint64_t num_sources, const uint32_t* sources, uint64_t cap, uint32_t* nodes,
                    double* probs, uint32_t* counts);
struct SearchArgs;
int launch_search_coop(rl_bh_ctx* c, const SearchArgs& sa);
int launch_level1(rl_bh_ctx* c, const SearchArgs& sa, const uint32_t* work, const uint64_t* nwork,
                  uint64_t run_below, uint32_t blocks, uint32_t chunk);
int launch_finalize(rl_bh_ctx* c, const SearchArgs& sa);
int launch_search_walk(rl_bh_ctx* c, const SearchArgs& sa);
int ensure_walk(rl_bh_ctx* c);   // group-walk buffers (FP32-mixed contexts), sized at init
// the context's level-1 search: the group walk (FP32-mixed by default), the cooperative wave per
// source, or one lane per source in the CPU path's order
inline bool uses_walk(const rl_bh_ctx* c) {
    return c->cfg.precision == RL_PREC_FP32_MIXED && (c->cfg.flags & (RL_FLAG_EXACT_ORDER | RL_FLAG_WAVE_PER_SOURCE)) == 0;
}
// Active sources from which an FP32-mixed update takes the group walk; below, one wavefront per source
// is faster.  A group's walk is a chain of about 10^4 dependent steps (7-9 ms alone on a wave), which
// pays off only when every wave has several groups.  Measured: one uniform network at the density of
// 16M neurons in a 25,000 um cube crosses between 0.87M and 1.7M active sources, while the updates of
// a 256-network simulation with 1.2M-2.3M searches still run 10 ms faster with a wavefront per
// source; from 3.5M on the walk is ahead in both.  Hence 192 groups of 64 per CU, 3,145,728 on 256
// CUs: the default of walk_min, which the autotuner moves (rl_bh_autotune).  RL_FLAG_WALK_ALWAYS: from 0.
inline uint64_t walk_min_default(const rl_bh_ctx* c) {
    return 64ULL * 192ULL * c->cus;
}
inline uint64_t walk_min_sources(const rl_bh_ctx* c) {
    return (c->cfg.flags & RL_FLAG_WALK_ALWAYS) != 0 ? 0 : c->walk_min;
}
int launch_descend_lane(rl_bh_ctx* c, const SearchArgs& sa, const uint4* items, const unsigned long long* count);

// commit (rl_commit.hip)
int enqueue_commit(rl_bh_ctx* c, const uint64_t* h_order);
int commit(rl_bh_ctx* c, uint32_t step, const uint64_t* h_order, uint64_t* created);

// neuron step (rl_step.hip)
int neurons_init(rl_bh_ctx* c, const rl_step_params* p, const double* x, const double* u, const double* ca,
                 const double* target_ca, const double* grown, const uint32_t* connected);
int neurons_step(rl_bh_ctx* c, uint32_t step, const double* input);
int enqueue_neuron_steps(rl_bh_ctx* c, uint32_t steps);  // steps d_steps[0] + k, k < steps
int enqueue_elements_commit(rl_bh_ctx* c);
int normalize_fired(rl_bh_ctx* c);      // the last fired status into the buffer a captured period reads
int neurons_commit(rl_bh_ctx* c, uint32_t update, uint32_t* deleted);
int neurons_get(rl_bh_ctx* c, double* x, double* u, double* ca, uint8_t* fired, double* delta, double* grown,
                uint32_t* connected);
int neurons_set_network_growth(rl_bh_ctx* c, const double* nu);
int neurons_set_element_params(rl_bh_ctx* c, const double* min_calcium, const double* retract_ratio);
void neurons_free(rl_bh_ctx* c);

// network graph (rl_net.hip)
int net_init(rl_bh_ctx* c, uint64_t capacity, const uint64_t* in_offsets, const uint32_t* sources,
             const int32_t* weights);
int net_get(rl_bh_ctx* c, uint64_t capacity, uint64_t* in_offsets, uint32_t* sources, int32_t* weights,
            uint64_t* num_edges);
int enqueue_net_append(rl_bh_ctx* c);   // appends the accepted requests of the last commit
int enqueue_net_delete(rl_bh_ctx* c, int type);  // deletes the synapses of the type's shrunken elements
int net_append(rl_bh_ctx* c);           // eager: grows the capacity if needed, then appends
int net_append_retry(rl_bh_ctx* c);     // after a captured append over capacity: grow, then append
int enqueue_syn_sync(rl_bh_ctx* c, int parity);  // out-edge lists from the graph, acc from fired_buf[parity]
int syn_prepare(rl_bh_ctx* c);          // enqueue_syn_sync at the current parity unless syn_valid
void net_free(rl_bh_ctx* c);

}  // namespace rl

#define RL_HIP(c, expr)                                   \
    do {                                                  \
        hipError_t rl_e_ = (expr);                        \
        if (rl_e_ != hipSuccess) {                        \
            return rl::hip_fail((c), rl_e_, #expr);       \
        }                                                 \
    } while (0)

#define RL_TRY(expr)                \
    do {                            \
        int rl_s_ = (expr);         \
        if (rl_s_ != RL_OK) {       \
            return rl_s_;           \
        }                           \
    } while (0)
