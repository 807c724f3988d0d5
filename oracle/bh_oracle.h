/*
 * CPU oracle for the single-rank Barnes-Hut connectivity update of RELeARN.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a restatement of the reference CPU
 * algorithm (pointer octree, stack traversal, sequential probability picker,
 * serial request commit, Izhikevich/calcium/growth step) used to check the HIP
 * path and to time the CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links it.
 *
 * All functions return ORC_OK (0) or a negative status.
 */
#ifndef RELEARN_BH_ORACLE_H
#define RELEARN_BH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_EINVAL = -1,     /* bad argument (sizes, theta range, sigma <= 0) */
    ORC_EOUTSIDE = -2,   /* position outside the simulation box */
    ORC_EDUPPOS = -3,    /* two neurons at the same position (reference loops forever) */
    ORC_EINTERNAL = -4,  /* an invariant of the tree was violated */
};

/* Draw streams for the target pick. */
enum {
    ORC_STREAM_COUNTER = 0, /* Philox4x32-10(seed; neuron, step, axon j, level) -> 32 bit draw */
    ORC_STREAM_MT19937 = 1, /* per OpenMP thread std::mt19937, reference seeding (parity unpinned) */
};

typedef struct orc_tree orc_tree;

/* a1/a2: pointer octree by serial insertion of neurons 0..n-1 (xyz is AoS n x 3). */
int orc_tree_build(uint64_t n, const double* xyz, const double box_min[3], const double box_max[3], orc_tree** out);
void orc_tree_free(orc_tree* t);
uint64_t orc_tree_num_nodes(const orc_tree* t);
uint32_t orc_tree_depth(const orc_tree* t);

/* a3/a4: leaf fill + upward pass.  enabled may be NULL (all enabled).  threads <= 0 means omp default. */
int orc_tree_update(orc_tree* t, const uint32_t* vac_exc, const uint32_t* vac_inh, const uint8_t* enabled, int threads);

/*
 * Canonical export in breadth-first order (level, then octant 0..7 below each parent), i.e. sorted by
 * (level, Morton prefix).  Arrays have orc_tree_num_nodes() entries (xyz arrays 3 per node), any may be NULL.
 *   level, nchild (0 for a leaf), link (first child index for inner nodes, neuron id for leaves),
 *   octant (octant of the node inside its parent, 0 for the root), box_lo/box_hi, side (max edge),
 *   cnt_exc/cnt_inh, has_exc/has_inh, pos_exc/pos_inh.
 */
int orc_tree_export(const orc_tree* t, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant,
                    double* box_lo, double* box_hi, double* side,
                    uint32_t* cnt_exc, uint32_t* cnt_inh, uint8_t* has_exc, uint8_t* has_inh,
                    double* pos_exc, double* pos_inh);

typedef struct {
    double theta;      /* acceptance criterion, 0 < theta <= 0.5 */
    double sigma;      /* Gaussian width, > 0 */
    double mu;         /* Gaussian offset */
    uint64_t seed;     /* draw-stream seed */
    uint32_t step;     /* plasticity-update counter (part of the counter-stream address) */
    int32_t stream;    /* ORC_STREAM_* */
    int32_t threads;   /* OpenMP threads, <= 0 means omp default */
    int32_t reserved;
} orc_search_params;

typedef struct {
    uint64_t searches;        /* find_target_neuron calls */
    uint64_t found;           /* searches that returned a target */
    uint64_t node_tests;      /* test_acceptance_criterion calls, all levels */
    uint64_t node_tests_l1;   /* of which in the level-1 (root) traversals */
    uint64_t candidates;      /* total length of all candidate lists */
    uint64_t picks;           /* pick_target calls that drew a number */
    uint64_t fallbacks;       /* candidate lists that needed the n/r fallback */
    uint64_t near_boundary;   /* searches with a draw within 1e-12*total of a cumulative boundary */
    double seconds;           /* wall time of the search phase */
} orc_search_stats;

/*
 * a5-a12: BarnesHut::find_target_neurons for neurons [first, first+count).  Every source neuron i gets
 * vac_axon[i] request slots starting at slot_offset[i - first]; out_target[slot] = target neuron id
 * for the j-th found request, out_count[i - first] = number of requests of neuron i.  Optional
 * per-slot diagnostics (may be NULL): out_margin[slot] = min over the levels of the search of the
 * distance of the draw to the nearest cumulative boundary divided by the total.  The tree must have
 * been updated with orc_tree_update.  signal: 0 excitatory, 1 inhibitory.
 */
int orc_find_targets(orc_tree* t, uint64_t n, const double* xyz, const uint8_t* signal,
                     const uint32_t* vac_axon, const uint8_t* enabled,
                     uint64_t first, uint64_t count, const uint64_t* slot_offset,
                     const orc_search_params* params,
                     uint32_t* out_target, uint32_t* out_count, double* out_margin,
                     uint32_t* out_node_tests, orc_search_stats* stats);

/*
 * One level of the search in isolation: candidate list of a traversal rooted at node `root` (BFS index)
 * for source position src (searching dendrites of signal sig), written as BFS node indices in list
 * order, with probabilities.  Returns the list length via out_len (capacity cap).  total gets the
 * sequential sum, fallback is set to 1 when the n/r fallback was used.
 */
int orc_candidate_list(const orc_tree* t, uint64_t root, const double src[3], uint32_t source_id, uint8_t sig,
                       double theta, double sigma, double mu, uint64_t cap,
                       uint64_t* out_nodes, double* out_prob, uint64_t* out_len, double* total, int* fallback,
                       uint64_t* node_tests);

/* ProbabilityPicker::pick_target(values, random_number); returns -1 on a reference exception. */
int64_t orc_pick_index(const double* prob, uint64_t len, double r);

/* Boost-style uniform real in [0, next) from one 32 bit draw (uniform_real_distribution on mt19937). */
double orc_draw_to_uniform(uint32_t k, double next_total);

/* Philox4x32-10 (Salmon et al., SC'11): out[4] = philox(ctr[4], key[2]). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* 32 bit counter-stream draw used for picks: tag 1. */
uint32_t orc_pick_draw(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t j, uint32_t level);

/* 64 bit request priority used for the commit order in counter mode: tag 2. */
uint64_t orc_commit_priority(uint64_t seed, uint32_t source, uint32_t step, uint32_t j);

/* std::mt19937 default seeded: k-th output (1-based). */
uint32_t orc_mt19937_kth(uint32_t seed, uint64_t k);

/* libstdc++ three argument std::hypot */
double orc_hypot3(double x, double y, double z);

/*
 * a14: ForwardConnector::process_requests + process_responses for a dense request list
 * (target, source, signal) of length m in the order the requests were appended.
 * order: permutation of 0..m-1 giving the processing order (the shuffled index order).  When NULL the
 * order is: counter mode -> sort by (orc_commit_priority(seed, source, step, j), index);
 * j is the per-source request ordinal (req_j).  vac_exc/vac_inh/vac_axon are updated in place,
 * conn_exc/conn_inh/conn_axon (may be NULL) incremented.  out_ok[i] = 1 when request i succeeded.
 */
int orc_commit(uint64_t m, const uint32_t* target, const uint32_t* source, const uint8_t* signal,
               const uint32_t* req_j, const uint64_t* order, uint64_t seed, uint32_t step,
               uint32_t* vac_exc, uint32_t* vac_inh, uint32_t* vac_axon,
               uint32_t* conn_exc, uint32_t* conn_inh, uint32_t* conn_axon, uint8_t* out_ok,
               uint64_t* out_created);

/* Fisher-Yates shuffle of shuffle.h with Boost uniform_int on std::mt19937 (parity unpinned). */
int orc_reference_shuffle(uint64_t m, uint32_t mt_seed, uint64_t* perm);

typedef struct {
    double a, b, c, d;        /* Izhikevich parameters */
    double v_spike;
    double k1, k2, k3;
    uint32_t h;               /* integration substeps */
    uint32_t reserved;
    double tau_ca, beta_ca;   /* calcium decay and increment per spike */
    double eta[3];            /* minimum calcium to grow: axon, exc dendrite, inh dendrite */
    double nu[3];             /* growth rates */
} orc_step_params;

/*
 * a15-a17: one simulation step for neurons 0..n-1: Izhikevich update with the given input,
 * calcium update, growth-curve deltas added to delta[3*n] (type-major: axon, exc, inh).
 * enabled may be NULL.  target_ca per neuron.  min_calcium: 3*n per-neuron minimum calcium
 * (type-major), or NULL for p->eta.
 */
int orc_neuron_step(uint64_t n, const orc_step_params* p, double* x, double* u, double* ca,
                    const double* input, uint8_t* fired, const double* target_ca,
                    double* delta, const uint8_t* enabled, const double* min_calcium);

/* a18: ElementBaseBase::update for one element type over n neurons; returns deletions per neuron. */
/* Synaptic input of the equally weighted model plus the background input: for every neuron the
 * weights (as doubles) of its in-edges whose source fired, summed in edge order, times conductance,
 * added to background[id] (NULL: 0).  In-edges are CSR by target (in_off n + 1, in_src, in_w). */
int orc_synaptic_input(uint64_t n, const uint64_t* in_off, const uint32_t* in_src, const int32_t* in_w,
                       const uint8_t* fired, double conductance, const double* background, double* out);

/* Network graph (NetworkGraph's local plastic edges, pairs merged, zero weights erased) and synapse
 * deletion of one element type (SynapseDeletionFinderCPU, counter-based draw). */
typedef struct orc_graph orc_graph;
int orc_graph_create(uint64_t n, const uint64_t* in_off, const uint32_t* src, const int32_t* w, orc_graph** out);
void orc_graph_free(orc_graph* g);
int orc_graph_add(orc_graph* g, uint64_t m, const uint32_t* target, const uint32_t* source, const int32_t* weight);
uint64_t orc_graph_num_edges(const orc_graph* g);
int orc_graph_export(const orc_graph* g, uint64_t cap, uint64_t* off, uint32_t* src, int32_t* w);
int orc_delete_synapses(orc_graph* g, int type, const uint32_t* to_delete, const uint32_t* net_of, const uint64_t* net_off,
                        const uint64_t* net_seed, uint32_t step, uint32_t* conn0, uint32_t* conn1, uint32_t* conn2,
                        uint32_t* vac0, uint32_t* vac1, uint32_t* vac2, const uint8_t* signal, uint64_t* deleted_total);
uint64_t orc_deletion_key(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t partner, uint32_t type, uint32_t copy);

int orc_elements_update(uint64_t n, double* grown, double* delta, uint32_t* connected, uint32_t* vacant,
                        const double* retract_ratio, uint32_t* out_deleted);

/* CPU baseline: omp parallel for + critical appends over the given source neurons (mt19937 stream). */
int orc_find_targets_faithful(orc_tree* t, uint64_t n, const double* xyz, const uint8_t* signal,
                              const uint32_t* vac_axon, const uint8_t* enabled,
                              const uint64_t* sources, uint64_t num_sources,
                              const orc_search_params* params, uint64_t* out_requests, orc_search_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
