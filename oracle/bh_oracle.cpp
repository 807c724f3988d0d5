// CPU oracle: restatement of RELeARN's single-rank Barnes-Hut connectivity update.
//
// TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker and the CPU baseline; the product library never links it.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference's relearn/ directory).  Built with g++ -O3 -fopenmp -ffp-contract=off, no
// -march: the reference sets no -march either, so its FP64 arithmetic is the same IEEE
// sequence of +,-,*,/ and the libstdc++ / glibc std::hypot, std::exp, std::sqrt.
//
// Parity pins (tests/test_oracle.py):
//   * octree (a1-a4): closed-form invariants of the capacity-1 PR octree, brute-force recursive
//     partition on small inputs, centroid recomputation in child order, the reference fixture
//     input/positions.txt (51,842 real positions);
//   * picker (a9): known answers of test/probability_picker/test_probability_picker.cpp:122-145 and
//     lower-bound equivalence (:80-100);
//   * Philox4x32-10: Random123 known-answer vectors, and the counter draws are uniform (KS, bit
//     balance, lag-1 correlation); std::mt19937: 10000th output 4123659995;
//   * Gaussian (a7), acceptance (a5): closed forms, monotonicity, brute-force list on tiny trees;
//   * commit (a14): the serial loop over the same order, written out in the test;
//   * step (a15-a17): closed-form single-substep recurrences, growth-curve zeros;
//   * synaptic input: a dense matrix product of the weight matrix with the fired vector;
//   * synapse deletion: conservation of units and element counts, all-or-nothing cases, a uniform
//     distribution of the deleted subset, and independence from the order the pairs are stored in.
// Parity unpinned: the reference's mt19937 seeding through boost::hash_combine (Boost 1.89.0),
// boost::random::uniform_int_distribution inside ::shuffle, the Boost normal input stream and the
// RandomHolder indices of RandomSynapseDeletionFinder; the first three only feed the mt19937
// CPU-baseline mode, and the deletion draw here is a counter-based uniform subset of the same size,
// so none of them is a parity check against the reference's own stream.

#include "bh_oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <numeric>
#include <random>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};

struct Box {
    V3 lo, hi;
};

// BoundingBox::contains (external/cpp-utility/data-structure/BoundingBox.hpp:84-90)
bool box_contains(const Box& b, const V3& p) {
    return b.lo.x <= p.x && p.x <= b.hi.x && b.lo.y <= p.y && p.y <= b.hi.y && b.lo.z <= p.z && p.z <= b.hi.z;
}

// Cell::get_octant_for_position (source/algorithm/Internal/octree/Cell.h:104-137)
unsigned octant_for(const Box& b, const V3& p) {
    unsigned idx = 0;
    idx |= (p.x < (b.lo.x + b.hi.x) / 2.0) ? 0U : 1U;
    idx |= (p.y < (b.lo.y + b.hi.y) / 2.0) ? 0U : 2U;
    idx |= (p.z < (b.lo.z + b.hi.z) / 2.0) ? 0U : 4U;
    return idx;
}

// Cell::get_size_for_octant (Cell.h:145-178); midpoint is std::midpoint per dimension
// (BoundingBox.hpp:144-146, Vec3.hpp:725-731).
Box box_for_octant(const Box& b, unsigned o) {
    const V3 mid{ std::midpoint(b.lo.x, b.hi.x), std::midpoint(b.lo.y, b.hi.y), std::midpoint(b.lo.z, b.hi.z) };
    Box r = b;
    if (o & 1U) {
        r.lo.x = mid.x;
    } else {
        r.hi.x = mid.x;
    }
    if (o & 2U) {
        r.lo.y = mid.y;
    } else {
        r.hi.y = mid.y;
    }
    if (o & 4U) {
        r.lo.z = mid.z;
    } else {
        r.hi.z = mid.z;
    }
    return r;
}

// BoundingBox::get_maximum_difference (BoundingBox.hpp:136-138) = max({dx, dy, dz}) (Vec3.hpp:653-655)
double max_edge(const Box& b) {
    return std::max({ b.hi.x - b.lo.x, b.hi.y - b.lo.y, b.hi.z - b.lo.z });
}

constexpr uint32_t kVirtual = 0xFFFFFFFFU;

// OctreeNode<BarnesHutCell> (OctreeNode.h:35-…, BarnesHutCell = BaseCell<true,true,false,false>,
// source/algorithm/BarnesHutInternal/BarnesHutCell.h:16): dendrite counts and positions for the two
// signal types, index 0 excitatory, 1 inhibitory.
struct Node {
    Node* child[8] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };
    Box box{};
    uint32_t neuron = kVirtual;
    uint32_t level = 0;
    bool parent = false;
    uint32_t cnt[2] = { 0, 0 };
    bool has_pos[2] = { false, false };
    V3 pos[2]{};
};

void set_neuron_position(Node* n, const V3& p) {
    // Cell::set_neuron_position (Cell.h:493-496) sets every element position of the cell
    n->pos[0] = p;
    n->pos[1] = p;
    n->has_pos[0] = true;
    n->has_pos[1] = true;
}

}  // namespace

struct orc_tree {
    Box box{};
    std::deque<Node> pool;
    Node* root = nullptr;
    std::vector<Node*> leaf_of;   // neuron id -> leaf (Octree::initializes_leaf_nodes)
    std::vector<Node*> bfs;       // canonical BFS order (lazily built)
    std::vector<uint64_t> bfs_index_of_leaf;
    uint32_t depth = 0;

    Node* alloc() {
        pool.emplace_back();
        return &pool.back();
    }

    void build_bfs() {
        bfs.clear();
        bfs.push_back(root);
        for (size_t i = 0; i < bfs.size(); i++) {
            Node* n = bfs[i];
            for (Node* c : n->child) {
                if (c != nullptr) {
                    bfs.push_back(c);
                }
            }
        }
    }
};

namespace {

// OctreeNode::insert (OctreeNode.h:153-250) with the 1-rank root of Octree::construct_global_tree_part
// (Octree.h:331-396): the root starts as a virtual leaf that the first real neuron replaces.
int tree_insert(orc_tree* t, const V3& p, uint32_t id) {
    if (!box_contains(t->root->box, p)) {
        return ORC_EOUTSIDE;
    }
    Node* parent_node = nullptr;
    for (Node* cur = t->root; cur != nullptr;) {
        const unsigned o = octant_for(cur->box, p);
        parent_node = cur;
        cur = cur->child[o];
    }
    if (!parent_node->parent) {
        if (parent_node->neuron == kVirtual && id != parent_node->neuron) {
            parent_node->neuron = id;
            set_neuron_position(parent_node, p);
            return ORC_OK;
        }
        for (;;) {
            const V3 pp = parent_node->pos[0];
            if (pp.x == p.x && pp.y == p.y && pp.z == p.z) {
                return ORC_EDUPPOS;  // the reference splits forever here
            }
            const unsigned own = octant_for(parent_node->box, pp);
            Node* nn = t->alloc();
            nn->box = box_for_octant(parent_node->box, own);
            set_neuron_position(nn, pp);
            nn->level = parent_node->level + 1;
            nn->neuron = parent_node->neuron;
            parent_node->child[own] = nn;
            parent_node->parent = true;
            parent_node->neuron = kVirtual;
            t->depth = std::max(t->depth, nn->level);
            if (octant_for(parent_node->box, p) == own) {
                parent_node = nn;
            } else {
                break;
            }
        }
    }
    const unsigned o = octant_for(parent_node->box, p);
    Node* nn = t->alloc();
    parent_node->child[o] = nn;
    parent_node->parent = true;
    nn->box = box_for_octant(parent_node->box, o);
    set_neuron_position(nn, p);
    nn->neuron = id;
    nn->level = parent_node->level + 1;
    t->depth = std::max(t->depth, nn->level);
    return ORC_OK;
}

// OctreeNodeUpdater::update_node (OctreeNodeHelper.h:43-152): for each dendrite type, sum the child
// counts in child order 0..7, accumulate position * (double)count, divide and clip to the cell box.
int update_node(Node* node) {
    for (int s = 0; s < 2; s++) {
        V3 acc{ 0.0, 0.0, 0.0 };
        uint32_t total = 0;
        for (Node* c : node->child) {
            if (c == nullptr) {
                continue;
            }
            const uint32_t n = c->cnt[s];
            total += n;
            if (!c->has_pos[s] && n != 0) {
                return ORC_EINTERNAL;
            }
            if (c->has_pos[s]) {
                if (!box_contains(c->box, c->pos[s])) {
                    return ORC_EINTERNAL;
                }
                const double w = static_cast<double>(n);
                acc.x += c->pos[s].x * w;
                acc.y += c->pos[s].y * w;
                acc.z += c->pos[s].z * w;
            }
        }
        node->cnt[s] = total;
        if (total == 0) {
            node->has_pos[s] = false;
        } else {
            const double w = static_cast<double>(total);
            V3 q{ acc.x / w, acc.y / w, acc.z / w };
            // Vec3::clip_to_box (Vec3.hpp:760-768)
            q.x = std::clamp(q.x, node->box.lo.x, node->box.hi.x);
            q.y = std::clamp(q.y, node->box.lo.y, node->box.hi.y);
            q.z = std::clamp(q.z, node->box.lo.z, node->box.hi.z);
            node->pos[s] = q;
            node->has_pos[s] = true;
        }
    }
    return ORC_OK;
}

// OctreeNodeUpdater::update_tree (OctreeNodeHelper.h:266-355): post-order with a visited flag.
int update_tree_serial(Node* tree) {
    struct Elem {
        Node* n;
        bool visited;
    };
    std::vector<Elem> stack;
    stack.push_back({ tree, false });
    while (!stack.empty()) {
        Elem& e = stack.back();
        if (e.visited) {
            if (e.n->parent) {
                if (int st = update_node(e.n); st != ORC_OK) {
                    return st;
                }
            }
            stack.pop_back();
            continue;
        }
        e.visited = true;
        Node* n = e.n;
        for (Node* c : n->child) {
            if (c != nullptr) {
                stack.push_back({ c, false });
            }
        }
    }
    return ORC_OK;
}

// Octree::update_tree_parallel (Octree.h:468-519): the subtrees two levels below the root are updated
// in parallel, then the upper part serially (children before parents).
int update_tree_parallel(Node* root, int threads) {
    std::vector<Node*> subtrees;
    std::vector<Node*> upper;
    upper.push_back(root);
    for (Node* c : root->child) {
        if (c == nullptr) {
            continue;
        }
        upper.push_back(c);
        for (Node* cc : c->child) {
            if (cc == nullptr) {
                continue;
            }
            upper.push_back(cc);
            subtrees.push_back(cc);
        }
    }
    int status = ORC_OK;
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < static_cast<int64_t>(subtrees.size()); i++) {
        const int st = update_tree_serial(subtrees[static_cast<size_t>(i)]);
        if (st != ORC_OK) {
#pragma omp critical(orc_status)
            status = st;
        }
    }
    if (status != ORC_OK) {
        return status;
    }
    while (!upper.empty()) {
        Node* n = upper.back();
        upper.pop_back();
        if (n->parent) {
            if (int st = update_node(n); st != ORC_OK) {
                return st;
            }
        }
    }
    return ORC_OK;
}

// ---------------------------------------------------------------- search

struct Stats {
    uint64_t node_tests = 0;
    uint64_t candidates = 0;
    uint64_t picks = 0;
    uint64_t fallbacks = 0;
};

enum class Acceptance { Discard, Expand, Accept };

// BarnesHutBase::test_acceptance_criterion (source/algorithm/BarnesHutInternal/BarnesHutBase.h:72-114)
Acceptance test_acceptance(const V3& src, const Node* node, int sig, double theta, Stats& st, double* dist_out) {
    st.node_tests++;
    if (node->cnt[sig] == 0) {
        return Acceptance::Discard;
    }
    const V3& tp = node->pos[sig];
    // Vec3::calculate_2_norm (Vec3.hpp:548-551) of target - source
    const double d = std::hypot(tp.x - src.x, tp.y - src.y, tp.z - src.z);
    *dist_out = d;
    if (d == 0.0) {
        return Acceptance::Discard;
    }
    if (!node->parent) {
        return Acceptance::Accept;
    }
    const double length = max_edge(node->box);
    return (length < (theta * d)) ? Acceptance::Accept : Acceptance::Expand;
}

// BarnesHutBase::get_nodes_to_consider (BarnesHutBase.h:129-206): LIFO stack, children pushed 0..7.
void nodes_to_consider(const V3& src, Node* root, int sig, double theta, Stats& st, std::vector<Node*>& out,
                       std::vector<Node*>& stack) {
    out.clear();
    if (root->cnt[sig] == 0) {
        return;
    }
    double d = 0.0;
    if (!root->parent) {
        if (test_acceptance(src, root, sig, theta, st, &d) != Acceptance::Discard) {
            out.push_back(root);
        }
        return;
    }
    stack.clear();
    for (Node* c : root->child) {
        if (c != nullptr) {
            stack.push_back(c);
        }
    }
    while (!stack.empty()) {
        Node* node = stack.back();
        stack.pop_back();
        const Acceptance a = test_acceptance(src, node, sig, theta, st, &d);
        if (a == Acceptance::Discard) {
            continue;
        }
        if (a == Acceptance::Accept) {
            out.push_back(node);
            continue;
        }
        for (Node* c : node->child) {
            if (c != nullptr) {
                stack.push_back(c);
            }
        }
    }
}

bool node_contains(const Node* n, uint32_t neuron) {
    // OctreeNode::contains (OctreeNode.h:79-91) at one rank
    return !n->parent && n->neuron == neuron;
}

// Kernel::create_probability_interval (source/algorithm/Kernel/Kernel.h:90-134) with the Gaussian
// kernel (Gaussian.h:47-67, KernelBase.h:49-54) and Kernel::calculate_attractiveness_to_connect (:61-75).
double probability_interval(const V3& src, uint32_t source_id, const std::vector<Node*>& nodes, int sig,
                            double sigma, double mu, std::vector<double>& prob, bool* fallback) {
    prob.clear();
    *fallback = false;
    if (nodes.empty()) {
        return 0.0;
    }
    const double scale_factor = -1.0 / (sigma * sigma);
    double sum = 0.0;
    for (const Node* n : nodes) {
        double p = 0.0;
        if (!node_contains(n, source_id)) {
            const V3& tp = n->pos[sig];
            const double d = std::hypot(tp.x - src.x, tp.y - src.y, tp.z - src.z);
            const double numerator = (d - mu) * (d - mu);
            const double exponent = numerator * scale_factor;
            const double exp_val = std::exp(exponent);
            p = (exp_val * 1.0) * static_cast<double>(n->cnt[sig]);
        }
        sum += p;
        prob.push_back(p);
    }
    if (sum == 0.0) {
        *fallback = true;
        prob.clear();
        for (const Node* n : nodes) {
            double p = 0.0;
            if (!node_contains(n, source_id)) {
                const V3& tp = n->pos[sig];
                p = static_cast<double>(n->cnt[sig]) / std::hypot(tp.x - src.x, tp.y - src.y, tp.z - src.z);
            }
            sum += p;
            prob.push_back(p);
        }
    }
    if (sum == 0.0) {
        prob.clear();
        return 0.0;
    }
    return sum;
}

// ProbabilityPicker::pick_target(values, random_number) (source/util/ProbabilityPicker.h:46-87).
// Returns -1 where the reference throws.
int64_t pick_index(const double* p, size_t len, double r, double* margin_out) {
    if (len == 0 || !(r >= 0.0)) {
        return -1;
    }
    for (size_t i = 0; i < len; i++) {
        if (!(p[i] >= 0.0)) {
            return -1;
        }
    }
    if (!(0.0 < r)) {
        for (size_t i = 0; i < len; i++) {
            if (p[i] > 0.0) {
                if (margin_out != nullptr) {
                    *margin_out = 0.0;
                }
                return static_cast<int64_t>(i);
            }
        }
        return -1;
    }
    size_t counter = 0;
    double s = 0.0;
    double before = 0.0;
    for (; counter < len && s < r; counter++) {
        before = s;
        s += p[counter];
    }
    if (!(s > 0.0)) {
        return -1;
    }
    if (margin_out != nullptr) {
        *margin_out = std::min(r - before, s - r);
    }
    while (p[counter - 1U] <= 0.0) {
        counter--;
    }
    return static_cast<int64_t>(counter - 1U);
}

// boost::random::uniform_real_distribution<double>(0, next) on a 32-bit engine: (k / 2^32) * next, retried
// while >= next (never happens for next > 0).
double draw_to_uniform(uint32_t k, double next) {
    const double numerator = static_cast<double>(k);
    const double divisor = 4294967295.0 + 1.0;
    return numerator / divisor * (next - 0.0) + 0.0;
}

// ---------------------------------------------------------------- Philox4x32-10

inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    const uint64_t prod = static_cast<uint64_t>(a) * static_cast<uint64_t>(b);
    *hi = static_cast<uint32_t>(prod >> 32U);
    *lo = static_cast<uint32_t>(prod);
}

void philox4x32_10(const uint32_t in[4], const uint32_t k[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = k[0], k1 = k[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B9U;
            k1 += 0xBB67AE85U;
        }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53U, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57U, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

uint32_t pick_draw(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t j, uint32_t level) {
    const uint32_t ctr[4] = { neuron, step, (j << 8U) | (level & 0xFFU), 1U };
    const uint32_t key[2] = { static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U) };
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return out[0];
}

uint64_t commit_priority(uint64_t seed, uint32_t source, uint32_t step, uint32_t j) {
    const uint32_t ctr[4] = { source, step, j, 2U };
    const uint32_t key[2] = { static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U) };
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return (static_cast<uint64_t>(out[0]) << 32U) | out[1];
}

// Deletion key of synapse unit `copy` between `neuron` and `partner` for element `type` (the same
// counter-based draw as the device; the reference draws indices with RandomHolder, parity unpinned).
uint64_t deletion_key(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t partner, uint32_t type, uint32_t copy) {
    const uint32_t ctr[4] = { neuron, step, partner, (copy << 4U) | (type << 2U) };
    const uint32_t key[2] = { static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U) };
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return (static_cast<uint64_t>(out[0]) << 32U) | out[1];
}

// boost::hash_combine of Boost >= 1.81 (64-bit size_t): seed = hash_mix(seed + 0x9e3779b9 + hash(v)).
// Used only by the mt19937 CPU-baseline stream; parity unpinned (Boost is not available here).
uint64_t boost_hash_mix(uint64_t x) {
    const uint64_t m = 0xe9846af9b1a615dULL;
    x ^= x >> 32U;
    x *= m;
    x ^= x >> 32U;
    x *= m;
    x ^= x >> 28U;
    return x;
}

uint64_t boost_hash_combine(uint64_t seed, uint64_t v) {
    return boost_hash_mix(seed + 0x9e3779b9ULL + v);
}

// RandomHolder::seed (source/util/Random.h:245-256) after relearn.cpp:817-829 at rank 0.
uint32_t reference_thread_seed(uint64_t user_seed, int thread_id) {
    uint64_t s = 0;
    s = boost_hash_combine(s, 0);
    s = boost_hash_combine(s, user_seed);
    return static_cast<uint32_t>(boost_hash_combine(s, static_cast<uint64_t>(thread_id)));
}

// boost::random::uniform_int_distribution on std::mt19937 for [0, range] (bucket rejection).
uint64_t boost_uniform_int(std::mt19937& gen, uint64_t range) {
    const uint64_t brange = 0xFFFFFFFFULL;
    if (range == 0) {
        return 0;
    }
    if (range == brange) {
        return gen();
    }
    if (range > brange) {
        // two-word composition; never reached for the request counts used here
        uint64_t r;
        do {
            r = (static_cast<uint64_t>(gen()) << 32U) | gen();
        } while (r > range && range != std::numeric_limits<uint64_t>::max());
        return r;
    }
    uint64_t bucket_size = brange / (range + 1);
    if (brange % (range + 1) == range) {
        ++bucket_size;
    }
    for (;;) {
        const uint64_t result = static_cast<uint64_t>(gen()) / bucket_size;
        if (result <= range) {
            return result;
        }
    }
}

// ::shuffle (source/util/shuffle/shuffle.h:63-79)
void reference_shuffle(std::vector<uint64_t>& v, std::mt19937& gen) {
    int64_t distance = static_cast<int64_t>(v.size());
    if (distance <= 1) {
        return;
    }
    size_t first = 0;
    size_t last = v.size() - 1;
    for (--distance; first < last; ++first, --distance) {
        const uint64_t rv = boost_uniform_int(gen, static_cast<uint64_t>(distance));
        if (rv != 0) {
            std::swap(v[first], v[first + rv]);
        }
    }
}

struct DrawSource {
    int mode = ORC_STREAM_COUNTER;
    uint64_t seed = 0;
    uint32_t step = 0;
    uint32_t neuron = 0;
    uint32_t j = 0;
    std::mt19937* mt = nullptr;

    uint32_t next(uint32_t level) {
        if (mode == ORC_STREAM_COUNTER) {
            return pick_draw(seed, neuron, step, j, level);
        }
        return (*mt)();
    }
};

struct SearchScratch {
    std::vector<Node*> list;
    std::vector<Node*> stack;
    std::vector<double> prob;
};

// BarnesHutBase::find_target_neuron (BarnesHutBase.h:223-255) + Kernel::pick_target (Kernel.h:147-173)
// + ProbabilityPicker::pick_target(values, key) (ProbabilityPicker.h:97-112).
// Returns the target neuron id or kVirtual when no target was found.
uint32_t find_target_neuron(Node* root, uint32_t source_id, const V3& src, int sig, double theta, double sigma,
                            double mu, DrawSource& draws, Stats& st, SearchScratch& sc, double* margin) {
    *margin = std::numeric_limits<double>::infinity();
    if (node_contains(root, source_id)) {
        return kVirtual;
    }
    if (!root->parent) {
        return root->neuron;
    }
    Node* sub = root;
    for (uint32_t level = 1;; level++) {
        nodes_to_consider(src, sub, sig, theta, st, sc.list, sc.stack);
        st.candidates += sc.list.size();
        if (sc.list.empty()) {
            return kVirtual;
        }
        bool fb = false;
        const double total = probability_interval(src, source_id, sc.list, sig, sigma, mu, sc.prob, &fb);
        if (fb) {
            st.fallbacks++;
        }
        if (total == 0.0) {
            return kVirtual;
        }
        double acc = 0.0;
        for (double p : sc.prob) {
            acc += p;  // ranges::accumulate
        }
        const double next = std::nextafter(acc, acc + 0.00001);
        double r;
        do {
            r = draw_to_uniform(draws.next(level), next);
        } while (!(r < next));
        st.picks++;
        double m = 0.0;
        const int64_t idx = pick_index(sc.prob.data(), sc.prob.size(), r, &m);
        if (idx < 0) {
            return kVirtual;
        }
        *margin = std::min(*margin, m / acc);
        Node* sel = sc.list[static_cast<size_t>(idx)];
        if (!sel->parent) {
            return sel->neuron;
        }
        sub = sel;
    }
}

int check_search_params(const orc_search_params* p) {
    if (p == nullptr) {
        return ORC_EINVAL;
    }
    // BarnesHutBase.h:74-77 (Constants::bh_max_theta = 0.5, source/Config.h:56-57), Gaussian.h:50
    if (!(p->theta > 0.0) || !(p->theta <= 0.5) || !(p->sigma > 0.0)) {
        return ORC_EINVAL;
    }
    if (p->stream != ORC_STREAM_COUNTER && p->stream != ORC_STREAM_MT19937) {
        return ORC_EINVAL;
    }
    return ORC_OK;
}

double now_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int orc_tree_build(uint64_t n, const double* xyz, const double box_min[3], const double box_max[3], orc_tree** out) {
    if (out == nullptr || box_min == nullptr || box_max == nullptr || (n > 0 && xyz == nullptr) || n >= kVirtual) {
        return ORC_EINVAL;
    }
    *out = nullptr;
    auto* t = new orc_tree();
    t->box = Box{ { box_min[0], box_min[1], box_min[2] }, { box_max[0], box_max[1], box_max[2] } };
    if (!(t->box.lo.x <= t->box.hi.x && t->box.lo.y <= t->box.hi.y && t->box.lo.z <= t->box.hi.z)) {
        delete t;
        return ORC_EINVAL;
    }
    // Octree::construct_global_tree_part at level 0: virtual root at the box centre
    t->root = t->alloc();
    t->root->box = t->box;
    t->root->level = 0;
    t->root->neuron = kVirtual;
    const V3 centre{ (0.0 * (box_max[0] - box_min[0])) + ((box_max[0] - box_min[0]) / 2.0) + box_min[0],
                     (0.0 * (box_max[1] - box_min[1])) + ((box_max[1] - box_min[1]) / 2.0) + box_min[1],
                     (0.0 * (box_max[2] - box_min[2])) + ((box_max[2] - box_min[2]) / 2.0) + box_min[2] };
    set_neuron_position(t->root, centre);
    // OctreeAlgorithm::init (source/algorithm/Internal/OctreeAlgorithm.h:58-67)
    for (uint64_t i = 0; i < n; i++) {
        const V3 p{ xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2] };
        const int st = tree_insert(t, p, static_cast<uint32_t>(i));
        if (st != ORC_OK) {
            delete t;
            return st;
        }
    }
    // Octree::initializes_leaf_nodes (Octree.h:230-270)
    t->leaf_of.assign(n, nullptr);
    std::vector<Node*> stack{ t->root };
    while (!stack.empty()) {
        Node* node = stack.back();
        stack.pop_back();
        if (!node->parent) {
            if (node->neuron == kVirtual) {
                continue;  // empty tree: virtual root
            }
            if (node->neuron >= n || t->leaf_of[node->neuron] != nullptr) {
                delete t;
                return ORC_EINTERNAL;
            }
            t->leaf_of[node->neuron] = node;
            continue;
        }
        for (Node* c : node->child) {
            if (c != nullptr) {
                stack.push_back(c);
            }
        }
    }
    for (uint64_t i = 0; i < n; i++) {
        if (t->leaf_of[i] == nullptr) {
            delete t;
            return ORC_EINTERNAL;
        }
    }
    t->build_bfs();
    *out = t;
    return ORC_OK;
}

void orc_tree_free(orc_tree* t) {
    delete t;
}

uint64_t orc_tree_num_nodes(const orc_tree* t) {
    return t == nullptr ? 0 : t->bfs.size();
}

uint32_t orc_tree_depth(const orc_tree* t) {
    return t == nullptr ? 0 : t->depth;
}

// OctreeAlgorithm::update_tree (OctreeAlgorithm.h:100-112) = update_leaf_nodes (:154-238) +
// Octree::synchronize_tree (Octree.h:275-281; the branch-node exchange is the identity at one rank).
int orc_tree_update(orc_tree* t, const uint32_t* vac_exc, const uint32_t* vac_inh, const uint8_t* enabled, int threads) {
    if (t == nullptr || (!t->leaf_of.empty() && (vac_exc == nullptr || vac_inh == nullptr))) {
        return ORC_EINVAL;
    }
    const size_t n = t->leaf_of.size();
    for (size_t i = 0; i < n; i++) {
        Node* leaf = t->leaf_of[i];
        if (enabled != nullptr && enabled[i] == 0) {
            leaf->cnt[0] = 0;
            leaf->cnt[1] = 0;
            continue;
        }
        leaf->cnt[0] = vac_exc[i];
        leaf->cnt[1] = vac_inh[i];
    }
    if (!t->root->parent) {
        return ORC_OK;
    }
    return update_tree_parallel(t->root, threads);
}

int orc_tree_export(const orc_tree* t, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant,
                    double* box_lo, double* box_hi, double* side, uint32_t* cnt_exc, uint32_t* cnt_inh,
                    uint8_t* has_exc, uint8_t* has_inh, double* pos_exc, double* pos_inh) {
    if (t == nullptr) {
        return ORC_EINVAL;
    }
    // BFS index of each node for the first-child links
    std::vector<uint64_t> first_child(t->bfs.size(), 0);
    {
        uint64_t next = 1;
        for (size_t i = 0; i < t->bfs.size(); i++) {
            first_child[i] = next;
            for (Node* c : t->bfs[i]->child) {
                if (c != nullptr) {
                    next++;
                }
            }
        }
    }
    for (size_t i = 0; i < t->bfs.size(); i++) {
        const Node* nd = t->bfs[i];
        uint8_t nc = 0;
        for (Node* c : nd->child) {
            if (c != nullptr) {
                nc++;
            }
        }
        if (octant != nullptr) {
            uint8_t k = 0;
            for (unsigned o = 0; o < 8; o++) {
                if (nd->child[o] != nullptr) {
                    octant[first_child[i] + k] = static_cast<uint8_t>(o);
                    k++;
                }
            }
            if (i == 0) {
                octant[0] = 0;
            }
        }
        if (level != nullptr) {
            level[i] = static_cast<uint8_t>(nd->level);
        }
        if (nchild != nullptr) {
            nchild[i] = nc;
        }
        if (link != nullptr) {
            link[i] = nd->parent ? static_cast<uint32_t>(first_child[i]) : nd->neuron;
        }
        if (box_lo != nullptr) {
            box_lo[3 * i] = nd->box.lo.x;
            box_lo[3 * i + 1] = nd->box.lo.y;
            box_lo[3 * i + 2] = nd->box.lo.z;
        }
        if (box_hi != nullptr) {
            box_hi[3 * i] = nd->box.hi.x;
            box_hi[3 * i + 1] = nd->box.hi.y;
            box_hi[3 * i + 2] = nd->box.hi.z;
        }
        if (side != nullptr) {
            side[i] = max_edge(nd->box);
        }
        if (cnt_exc != nullptr) {
            cnt_exc[i] = nd->cnt[0];
        }
        if (cnt_inh != nullptr) {
            cnt_inh[i] = nd->cnt[1];
        }
        if (has_exc != nullptr) {
            has_exc[i] = nd->has_pos[0] ? 1 : 0;
        }
        if (has_inh != nullptr) {
            has_inh[i] = nd->has_pos[1] ? 1 : 0;
        }
        if (pos_exc != nullptr) {
            pos_exc[3 * i] = nd->has_pos[0] ? nd->pos[0].x : 0.0;
            pos_exc[3 * i + 1] = nd->has_pos[0] ? nd->pos[0].y : 0.0;
            pos_exc[3 * i + 2] = nd->has_pos[0] ? nd->pos[0].z : 0.0;
        }
        if (pos_inh != nullptr) {
            pos_inh[3 * i] = nd->has_pos[1] ? nd->pos[1].x : 0.0;
            pos_inh[3 * i + 1] = nd->has_pos[1] ? nd->pos[1].y : 0.0;
            pos_inh[3 * i + 2] = nd->has_pos[1] ? nd->pos[1].z : 0.0;
        }
    }
    return ORC_OK;
}

int orc_find_targets(orc_tree* t, uint64_t n, const double* xyz, const uint8_t* signal, const uint32_t* vac_axon,
                     const uint8_t* enabled, uint64_t first, uint64_t count, const uint64_t* slot_offset,
                     const orc_search_params* params, uint32_t* out_target, uint32_t* out_count, double* out_margin,
                     uint32_t* out_node_tests, orc_search_stats* stats) {
    if (t == nullptr || xyz == nullptr || signal == nullptr || vac_axon == nullptr || slot_offset == nullptr ||
        out_target == nullptr || out_count == nullptr || n != t->leaf_of.size() || first + count > n) {
        return ORC_EINVAL;
    }
    if (int st = check_search_params(params); st != ORC_OK) {
        return st;
    }
    const double t0 = now_seconds();
    uint64_t searches = 0, found = 0, tests = 0, tests_l1 = 0, cands = 0, picks = 0, fbs = 0, nearb = 0;
    const int nt = params->threads > 0 ? params->threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(+ : searches, found, tests, tests_l1, cands, picks, fbs, nearb)
    {
        std::mt19937 mt(reference_thread_seed(params->seed, omp_get_thread_num()));
        SearchScratch sc;
        DrawSource ds;
        ds.mode = params->stream;
        ds.seed = params->seed;
        ds.step = params->step;
        ds.mt = &mt;
#pragma omp for schedule(static)
        for (int64_t ii = 0; ii < static_cast<int64_t>(count); ii++) {
            const uint64_t i = first + static_cast<uint64_t>(ii);
            const uint64_t slot0 = slot_offset[ii];
            out_count[ii] = 0;
            // BarnesHut::find_target_neurons (source/algorithm/BarnesHutInternal/BarnesHut.cpp:57-86)
            if (enabled != nullptr && enabled[i] == 0) {
                continue;
            }
            const uint32_t na = vac_axon[i];
            if (na == 0) {
                continue;
            }
            const V3 src{ xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2] };
            const int sig = signal[i] == 0 ? 0 : 1;
            ds.neuron = static_cast<uint32_t>(i);
            // BarnesHutBase::find_target_neurons (BarnesHutBase.h:272-298)
            for (uint32_t j = 0; j < na; j++) {
                ds.j = j;
                Stats st;
                double margin = 0.0;
                const uint32_t tgt =
                    find_target_neuron(t->root, static_cast<uint32_t>(i), src, sig, params->theta, params->sigma, params->mu,
                                       ds, st, sc, &margin);
                searches++;
                tests += st.node_tests;
                cands += st.candidates;
                picks += st.picks;
                fbs += st.fallbacks;
                if (margin < 1e-12) {
                    nearb++;
                }
                if (out_node_tests != nullptr) {
                    out_node_tests[slot0 + j] = static_cast<uint32_t>(std::min<uint64_t>(st.node_tests, 0xFFFFFFFFULL));
                }
                if (out_margin != nullptr) {
                    out_margin[slot0 + j] = margin;
                }
                if (tgt == kVirtual) {
                    break;
                }
                found++;
                out_target[slot0 + j] = tgt;
                out_count[ii] = j + 1;
            }
        }
    }
    if (stats != nullptr) {
        stats->searches = searches;
        stats->found = found;
        stats->node_tests = tests;
        stats->node_tests_l1 = tests_l1;
        stats->candidates = cands;
        stats->picks = picks;
        stats->fallbacks = fbs;
        stats->near_boundary = nearb;
        stats->seconds = now_seconds() - t0;
    }
    return ORC_OK;
}

int orc_find_targets_faithful(orc_tree* t, uint64_t n, const double* xyz, const uint8_t* signal, const uint32_t* vac_axon,
                              const uint8_t* enabled, const uint64_t* sources, uint64_t num_sources,
                              const orc_search_params* params, uint64_t* out_requests, orc_search_stats* stats) {
    if (t == nullptr || xyz == nullptr || signal == nullptr || vac_axon == nullptr || sources == nullptr ||
        n != t->leaf_of.size()) {
        return ORC_EINVAL;
    }
    if (int st = check_search_params(params); st != ORC_OK) {
        return st;
    }
    for (uint64_t k = 0; k < num_sources; k++) {
        if (sources[k] >= n) {
            return ORC_EINVAL;
        }
    }
    struct Request {
        uint32_t target, source;
        uint8_t signal;
    };
    std::vector<Request> requests;  // CommunicationVector of the single rank
    const double t0 = now_seconds();
    uint64_t searches = 0, found = 0, tests = 0, cands = 0, picks = 0, fbs = 0;
    const int nt = params->threads > 0 ? params->threads : omp_get_max_threads();
    // BarnesHut::find_target_neurons (BarnesHut.cpp:57-86): static schedule, critical appends, a
    // per-thread std::mt19937 for the Algorithm key.
#pragma omp parallel num_threads(nt) reduction(+ : searches, found, tests, cands, picks, fbs)
    {
        std::mt19937 mt(reference_thread_seed(params->seed, omp_get_thread_num()));
        DrawSource ds;
        ds.mode = params->stream;
        ds.seed = params->seed;
        ds.step = params->step;
        ds.mt = &mt;
#pragma omp for schedule(static)
        for (int64_t k = 0; k < static_cast<int64_t>(num_sources); k++) {
            const uint64_t i = sources[k];
            if (enabled != nullptr && enabled[i] == 0) {
                continue;
            }
            const uint32_t na = vac_axon[i];
            if (na == 0) {
                continue;
            }
            const V3 src{ xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2] };
            const int sig = signal[i] == 0 ? 0 : 1;
            ds.neuron = static_cast<uint32_t>(i);
            // per-search allocations as in the reference (fresh vectors inside each call)
            std::vector<Request> local;
            local.reserve(na);
            for (uint32_t j = 0; j < na; j++) {
                ds.j = j;
                SearchScratch sc;
                sc.list.reserve(30);
                sc.stack.reserve(30);
                Stats st;
                double margin = 0.0;
                const uint32_t tgt = find_target_neuron(t->root, static_cast<uint32_t>(i), src, sig, params->theta,
                                                        params->sigma, params->mu, ds, st, sc, &margin);
                searches++;
                tests += st.node_tests;
                cands += st.candidates;
                picks += st.picks;
                fbs += st.fallbacks;
                if (tgt == kVirtual) {
                    break;
                }
                found++;
                local.push_back({ tgt, static_cast<uint32_t>(i), static_cast<uint8_t>(sig) });
            }
            for (const Request& r : local) {
#pragma omp critical(BHrequests)
                requests.push_back(r);
            }
        }
    }
    if (out_requests != nullptr) {
        *out_requests = requests.size();
    }
    if (stats != nullptr) {
        stats->searches = searches;
        stats->found = found;
        stats->node_tests = tests;
        stats->node_tests_l1 = 0;
        stats->candidates = cands;
        stats->picks = picks;
        stats->fallbacks = fbs;
        stats->near_boundary = 0;
        stats->seconds = now_seconds() - t0;
    }
    return ORC_OK;
}

int orc_candidate_list(const orc_tree* t, uint64_t root, const double src[3], uint32_t source_id, uint8_t sig,
                       double theta, double sigma, double mu, uint64_t cap, uint64_t* out_nodes, double* out_prob,
                       uint64_t* out_len, double* total, int* fallback, uint64_t* node_tests) {
    if (t == nullptr || src == nullptr || root >= t->bfs.size() || out_len == nullptr || !(theta > 0.0) ||
        !(theta <= 0.5) || !(sigma > 0.0)) {
        return ORC_EINVAL;
    }
    std::vector<uint64_t> index_of;
    // map node pointer -> BFS index through a sorted vector of pointers
    std::vector<std::pair<const Node*, uint64_t>> idx;
    idx.reserve(t->bfs.size());
    for (uint64_t i = 0; i < t->bfs.size(); i++) {
        idx.emplace_back(t->bfs[i], i);
    }
    std::sort(idx.begin(), idx.end());
    Stats st;
    SearchScratch sc;
    const V3 s{ src[0], src[1], src[2] };
    nodes_to_consider(s, t->bfs[root], sig == 0 ? 0 : 1, theta, st, sc.list, sc.stack);
    bool fb = false;
    const double sum = probability_interval(s, source_id, sc.list, sig == 0 ? 0 : 1, sigma, mu, sc.prob, &fb);
    *out_len = sc.list.size();
    if (total != nullptr) {
        *total = sum;
    }
    if (fallback != nullptr) {
        *fallback = fb ? 1 : 0;
    }
    if (node_tests != nullptr) {
        *node_tests = st.node_tests;
    }
    for (uint64_t k = 0; k < sc.list.size() && k < cap; k++) {
        if (out_nodes != nullptr) {
            const auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(static_cast<const Node*>(sc.list[k]), uint64_t{ 0 }));
            out_nodes[k] = it->second;
        }
        if (out_prob != nullptr) {
            out_prob[k] = k < sc.prob.size() ? sc.prob[k] : 0.0;
        }
    }
    return ORC_OK;
}

int64_t orc_pick_index(const double* prob, uint64_t len, double r) {
    if (prob == nullptr && len > 0) {
        return -1;
    }
    return pick_index(prob, static_cast<size_t>(len), r, nullptr);
}

double orc_draw_to_uniform(uint32_t k, double next_total) {
    return draw_to_uniform(k, next_total);
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr, key, out);
}

uint32_t orc_pick_draw(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t j, uint32_t level) {
    return pick_draw(seed, neuron, step, j, level);
}

uint64_t orc_deletion_key(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t partner, uint32_t type, uint32_t copy) {
    return deletion_key(seed, neuron, step, partner, type, copy);
}

uint64_t orc_commit_priority(uint64_t seed, uint32_t source, uint32_t step, uint32_t j) {
    return commit_priority(seed, source, step, j);
}

uint32_t orc_mt19937_kth(uint32_t seed, uint64_t k) {
    std::mt19937 g(seed);
    uint32_t v = 0;
    for (uint64_t i = 0; i < k; i++) {
        v = g();
    }
    return v;
}

double orc_hypot3(double x, double y, double z) {
    return std::hypot(x, y, z);
}

// ForwardConnector::process_requests (source/algorithm/Connector.cpp:33-121) and process_responses
// (:123-197) at one rank; ElementBaseBase::connect_elements (neurons/synaptic_elements/ElementBaseBase.h:274-283).
int orc_commit(uint64_t m, const uint32_t* target, const uint32_t* source, const uint8_t* signal, const uint32_t* req_j,
               const uint64_t* order, uint64_t seed, uint32_t step, uint32_t* vac_exc, uint32_t* vac_inh,
               uint32_t* vac_axon, uint32_t* conn_exc, uint32_t* conn_inh, uint32_t* conn_axon, uint8_t* out_ok,
               uint64_t* out_created) {
    if (m > 0 && (target == nullptr || source == nullptr || signal == nullptr || vac_exc == nullptr || vac_inh == nullptr ||
                  vac_axon == nullptr || out_ok == nullptr)) {
        return ORC_EINVAL;
    }
    if (order == nullptr && m > 0 && req_j == nullptr) {
        return ORC_EINVAL;
    }
    std::vector<uint64_t> perm(m);
    if (order != nullptr) {
        std::vector<uint8_t> seen(m, 0);
        for (uint64_t k = 0; k < m; k++) {
            if (order[k] >= m || seen[order[k]] != 0) {
                return ORC_EINVAL;
            }
            seen[order[k]] = 1;
            perm[k] = order[k];
        }
    } else {
        std::vector<std::pair<uint64_t, uint64_t>> keyed(m);
        for (uint64_t k = 0; k < m; k++) {
            keyed[k] = { commit_priority(seed, source[k], step, req_j[k]), k };
        }
        std::sort(keyed.begin(), keyed.end());
        for (uint64_t k = 0; k < m; k++) {
            perm[k] = keyed[k].second;
        }
    }
    uint64_t created = 0;
    for (uint64_t k = 0; k < m; k++) {
        out_ok[k] = 0;
    }
    for (uint64_t k = 0; k < m; k++) {
        const uint64_t i = perm[k];
        const uint32_t tgt = target[i];
        if (tgt == source[i]) {
            continue;
        }
        uint32_t* vac = signal[i] == 0 ? vac_exc : vac_inh;
        uint32_t* conn = signal[i] == 0 ? conn_exc : conn_inh;
        if (vac[tgt] == 0) {
            continue;
        }
        vac[tgt]--;
        if (conn != nullptr) {
            conn[tgt]++;
        }
        out_ok[i] = 1;
        created++;
    }
    for (uint64_t i = 0; i < m; i++) {
        if (out_ok[i] == 0) {
            continue;
        }
        const uint32_t s = source[i];
        if (vac_axon[s] == 0) {
            return ORC_EINTERNAL;  // process_responses: source had no vacant axon
        }
        vac_axon[s]--;
        if (conn_axon != nullptr) {
            conn_axon[s]++;
        }
    }
    if (out_created != nullptr) {
        *out_created = created;
    }
    return ORC_OK;
}

int orc_reference_shuffle(uint64_t m, uint32_t mt_seed, uint64_t* perm) {
    if (m > 0 && perm == nullptr) {
        return ORC_EINVAL;
    }
    std::vector<uint64_t> v(m);
    std::iota(v.begin(), v.end(), 0);
    std::mt19937 g(mt_seed);
    reference_shuffle(v, g);
    for (uint64_t k = 0; k < m; k++) {
        perm[k] = v[k];
    }
    return ORC_OK;
}

// gaussian_growth_curve (source/neurons/synaptic_elements/GrowthCurves.h:30-54)
static double gaussian_growth_curve(double current, double eta, double epsilon) {
    if (eta == epsilon) {
        if (current == eta) {
            return 0.0;
        }
        return -1.0;
    }
    constexpr double factor = 1.6651092223153955127063292897904020952611777045288814583336582344;
    constexpr double factor_inv = 1.0 / factor;
    const double xi = (eta + epsilon) * 0.5;
    const double zeta = (eta - epsilon) * factor_inv;
    const double difference = current - xi;
    const double quotient = difference / zeta;
    const double product = quotient * quotient;
    return (2.0 * std::exp(-product)) - 1.0;
}

int orc_neuron_step(uint64_t n, const orc_step_params* p, double* x, double* u, double* ca, const double* input,
                    uint8_t* fired, const double* target_ca, double* delta, const uint8_t* enabled,
                    const double* min_calcium) {
    if (p == nullptr || (n > 0 && (x == nullptr || u == nullptr || ca == nullptr || input == nullptr || fired == nullptr ||
                                   target_ca == nullptr || delta == nullptr)) ||
        p->h == 0) {
        return ORC_EINVAL;
    }
    // IzhikevichModelCPU::update_activity (source/neurons/models/izhikevich/IzhikevichModelCPU.cpp:75-125)
    const double scale = 1.0 / static_cast<double>(p->h);
    const double ab = p->a * p->b;
    for (uint64_t i = 0; i < n; i++) {
        if (enabled != nullptr && enabled[i] == 0) {
            continue;
        }
        double xv = x[i];
        double uv = u[i];
        uint8_t spiked = 0;
        for (uint32_t s = 0; s < p->h; s++) {
            const double x_increase = (p->k1 * xv * xv) + (p->k2 * xv) + p->k3 - uv + input[i];
            const double u_increase = (ab * xv) - (p->a * uv);
            xv += x_increase * scale;
            uv += u_increase * scale;
            if (xv >= p->v_spike) {
                xv = p->c;
                uv += p->d;
                spiked = 1;
                break;
            }
        }
        fired[i] = spiked;
        x[i] = xv;
        u[i] = uv;
    }
    // CalciumCalculatorCPU::update_current_calcium (source/neurons/calcium/CalciumCalculatorCPU.cpp:40-103)
    const double tau_inv = -1.0 / p->tau_ca;
    for (uint64_t i = 0; i < n; i++) {
        if (enabled != nullptr && enabled[i] == 0) {
            continue;
        }
        double c = ca[i];
        const double local_beta = p->beta_ca * static_cast<double>(fired[i] == 1);
        for (uint32_t s = 0; s < p->h; s++) {
            c += scale * (c * tau_inv + local_beta);
        }
        ca[i] = c;
    }
    // update_number_elements_openmp (source/neurons/synaptic_elements/SynapticElementsCPU.cpp:25-49) for the
    // axon, excitatory and inhibitory dendrite types (:73-75), added with add_to_delta.  The minimum
    // calcium is per neuron (get_minimum_calcium, :62; the --synaptic-elements-file values) or one value
    // per type (the constant calculators of relearn.cpp:1199-1201).
    for (int ty = 0; ty < 3; ty++) {
        for (uint64_t i = 0; i < n; i++) {
            double inc = 0.0;
            if (enabled == nullptr || enabled[i] != 0) {
                const double eta = min_calcium != nullptr ? min_calcium[static_cast<uint64_t>(ty) * n + i] : p->eta[ty];
                const double clamped_target = std::max(target_ca[i], eta);
                const double dz = gaussian_growth_curve(ca[i], eta, clamped_target);
                inc = dz * p->nu[ty];
            }
            delta[static_cast<uint64_t>(ty) * n + i] += inc;
        }
    }
    return ORC_OK;
}

// SynapticEquallyWeightedActivityInputCPU::update_local_input
// (source/neurons/input/SynapticEquallyWeightedActivityInputCPU.cpp:21-50): per neuron, the weights of
// the in-edges whose source fired, accumulated as doubles in edge order, times the conductance; then
// CombinedActivityInputCPU::update_input_range (source/neurons/input/CombinedActivityInputCPU.h:46-62)
// fills the input with 0 and adds the background input and the synaptic input in turn.
int orc_synaptic_input(uint64_t n, const uint64_t* in_off, const uint32_t* in_src, const int32_t* in_w,
                       const uint8_t* fired, double conductance, const double* background, double* out) {
    if (n > 0 && (in_off == nullptr || fired == nullptr || out == nullptr)) {
        return ORC_EINVAL;
    }
    for (uint64_t id = 0; id < n; id++) {
        double inp = 0.0;
        for (uint64_t e = in_off[id]; e < in_off[id + 1]; e++) {
            if (fired[in_src[e]] == 0) {
                continue;
            }
            inp += static_cast<double>(in_w[e]);
        }
        const double synaptic = inp * conductance;
        double total = 0.0;
        if (background != nullptr) {
            total += background[id];
        }
        total += synaptic;
        out[id] = total;
    }
    return ORC_OK;
}

// Local plastic part of NetworkGraph (source/neurons/NetworkGraph.h): per neuron its in-edges
// (source -> weight) and out-edges (target -> weight); add_synapse adds the weight to the pair and
// erases a pair whose weight becomes 0.
struct orc_graph {
    uint64_t n = 0;
    std::vector<std::map<uint32_t, int32_t>> in, out;

    void add(uint32_t target, uint32_t source, int32_t weight) {
        auto add_one = [](std::map<uint32_t, int32_t>& m, uint32_t key, int32_t w) {
            const int32_t now = (m[key] += w);
            if (now == 0) {
                m.erase(key);
            }
        };
        add_one(in[target], source, weight);
        add_one(out[source], target, weight);
    }
};

int orc_graph_create(uint64_t n, const uint64_t* in_off, const uint32_t* src, const int32_t* w, orc_graph** out) {
    if (out == nullptr) {
        return ORC_EINVAL;
    }
    auto* g = new orc_graph();
    g->n = n;
    g->in.resize(n);
    g->out.resize(n);
    if (in_off != nullptr) {
        for (uint64_t t = 0; t < n; t++) {
            for (uint64_t e = in_off[t]; e < in_off[t + 1]; e++) {
                if (src[e] >= n) {
                    delete g;
                    return ORC_EINVAL;
                }
                if (w[e] != 0) {
                    g->add(static_cast<uint32_t>(t), src[e], w[e]);
                }
            }
        }
    }
    *out = g;
    return ORC_OK;
}

void orc_graph_free(orc_graph* g) {
    delete g;
}

int orc_graph_add(orc_graph* g, uint64_t m, const uint32_t* target, const uint32_t* source, const int32_t* weight) {
    for (uint64_t i = 0; i < m; i++) {
        if (target[i] >= g->n || source[i] >= g->n) {
            return ORC_EINVAL;
        }
        g->add(target[i], source[i], weight[i]);
    }
    return ORC_OK;
}

uint64_t orc_graph_num_edges(const orc_graph* g) {
    uint64_t e = 0;
    for (const auto& m : g->in) {
        e += m.size();
    }
    return e;
}

int orc_graph_export(const orc_graph* g, uint64_t cap, uint64_t* off, uint32_t* src, int32_t* w) {
    uint64_t e = 0;
    off[0] = 0;
    for (uint64_t t = 0; t < g->n; t++) {
        for (const auto& [s, wt] : g->in[t]) {
            if (e >= cap) {
                return ORC_EINVAL;
            }
            src[e] = s;
            w[e] = wt;
            e++;
        }
        off[t + 1] = e;
    }
    return ORC_OK;
}

// SynapseDeletionFinderCPU::find_synapses_to_delete + commit_deletions
// (source/neurons/helper/SynapseDeletionFinderCPU.cpp) for one element type, single rank: every
// neuron with to_delete[i] > 0 lists its synapse units (SynapseDeletionFinderBase::register_synapses,
// SynapseDeletionFinderBase.cpp:33-151: axons its out-edges, dendrites its in-edges of the dendrite's
// sign, |weight| units per pair), draws to_delete[i] of them without replacement
// (RandomSynapseDeletionFinder::find_synapses_on_neuron, SynapseDeletionFinder.cpp:53-69; here the
// units with the smallest deletion_key), and all draws are made before any is committed.  Each
// committed unit lowers its pair's weight by one and disconnects one element of the partner
// (ElementBaseBase::disconnect_elements, ElementBaseBase.h:311-319).  Neuron ids inside the draw are
// local to the neuron's network.
int orc_delete_synapses(orc_graph* g, int type, const uint32_t* to_delete, const uint32_t* net_of, const uint64_t* net_off,
                        const uint64_t* net_seed, uint32_t step, uint32_t* conn0, uint32_t* conn1, uint32_t* conn2,
                        uint32_t* vac0, uint32_t* vac1, uint32_t* vac2, const uint8_t* signal, uint64_t* deleted_total) {
    if (g == nullptr || type < 0 || type > 2) {
        return ORC_EINVAL;
    }
    uint32_t* conn[3] = { conn0, conn1, conn2 };
    uint32_t* vac[3] = { vac0, vac1, vac2 };
    struct Request {
        uint32_t initiator, partner;
    };
    std::vector<Request> requests;
    auto local = [&](uint32_t i) { return i - static_cast<uint32_t>(net_off[net_of[i]]); };
    for (uint64_t i = 0; i < g->n; i++) {
        const uint32_t want = to_delete[i];
        if (want == 0) {
            continue;
        }
        struct U {
            uint64_t key;
            uint32_t partner_local, copy, partner;
        };
        std::vector<U> units;
        const auto& edges = type == 0 ? g->out[i] : g->in[i];
        for (const auto& [p, w] : edges) {
            if (type == 1 && w < 0) continue;   // searching excitatory synapses, found an inhibitory one
            if (type == 2 && w > 0) continue;
            const uint32_t aw = static_cast<uint32_t>(w < 0 ? -w : w);
            for (uint32_t c = 0; c < aw; c++) {
                units.push_back({ deletion_key(net_seed[net_of[i]], local(static_cast<uint32_t>(i)), step, local(p),
                                               static_cast<uint32_t>(type), c), local(p), c, p });
            }
        }
        if (want > units.size()) {
            return ORC_EINTERNAL;  // the reference throws: more deletions than synapses
        }
        std::sort(units.begin(), units.end(), [](const U& a, const U& b) {
            if (a.key != b.key) return a.key < b.key;
            if (a.partner_local != b.partner_local) return a.partner_local < b.partner_local;
            return a.copy < b.copy;
        });
        for (uint32_t k = 0; k < want; k++) {
            requests.push_back({ static_cast<uint32_t>(i), units[k].partner });
        }
    }
    for (const auto& r : requests) {
        if (type == 0) {
            const int32_t weight = signal[r.initiator] == 0 ? -1 : 1;   // excitatory pair loses +1
            g->add(r.partner, r.initiator, weight);
            const int dt = signal[r.initiator] == 0 ? 1 : 2;
            conn[dt][r.partner] -= 1;
            vac[dt][r.partner] += 1;
        } else {
            const int32_t weight = type == 1 ? -1 : 1;
            g->add(r.initiator, r.partner, weight);
            conn[0][r.partner] -= 1;
            vac[0][r.partner] += 1;
        }
    }
    if (deleted_total != nullptr) {
        *deleted_total = requests.size();
    }
    return ORC_OK;
}

// ElementBaseBase::update (source/neurons/synaptic_elements/ElementBaseBase.h:420-487)
int orc_elements_update(uint64_t n, double* grown, double* delta, uint32_t* connected, uint32_t* vacant,
                        const double* retract_ratio, uint32_t* out_deleted) {
    if (n > 0 && (grown == nullptr || delta == nullptr || connected == nullptr || vacant == nullptr || out_deleted == nullptr)) {
        return ORC_EINVAL;
    }
    for (uint64_t i = 0; i < n; i++) {
        const double current_count = grown[i];
        const double current_delta = delta[i];
        const uint32_t connected_integral = connected[i];
        const double retract = retract_ratio != nullptr ? retract_ratio[i] : 0.0;
        const double current_connected = static_cast<double>(connected_integral);
        const double current_vacant = current_count - current_connected;
        if (!(current_count >= 0.0) || !(current_vacant >= 0.0)) {
            return ORC_EINTERNAL;
        }
        if (const double new_vacant = current_vacant + current_delta; new_vacant >= 0.0) {
            const double new_vacant_after_retract = (1 - retract) * new_vacant;
            const double new_count = new_vacant_after_retract + current_connected;
            grown[i] = new_count;
            delta[i] = 0.0;
            vacant[i] = static_cast<uint32_t>(new_vacant_after_retract);
            out_deleted[i] = 0;
            continue;
        }
        if (current_count + current_delta <= 0.0) {
            grown[i] = 0.0;
            delta[i] = 0.0;
            vacant[i] = 0;
            connected[i] = 0;
            out_deleted[i] = connected_integral;
            continue;
        }
        const double new_count = current_count + current_delta;
        const double new_connected_count = std::floor(new_count);
        const double num_vacant = new_count - new_connected_count;
        const double retracted_new_count = ((1 - retract) * num_vacant) + new_connected_count;
        grown[i] = retracted_new_count;
        delta[i] = 0.0;
        vacant[i] = 0;
        connected[i] = static_cast<uint32_t>(new_connected_count);
        out_deleted[i] = static_cast<uint32_t>(current_connected - new_connected_count);
    }
    return ORC_OK;
}

}  // extern "C"
