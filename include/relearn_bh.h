/*
 * relearn_bh.h -- C ABI of the MI355X (gfx950) Barnes-Hut connectivity update.
 *
 * Drop-in boundary for RELeARN's `--algorithm barnes-hut` on one rank.  Each entry point replaces
 * a member of the reference's Algorithm interface (source/algorithm/Algorithm.h:48-197, paths relative
 * to the reference's relearn/ directory) as implemented by BarnesHut
 * (source/algorithm/BarnesHutInternal/BarnesHut.{h,cpp}) on top of OctreeAlgorithm and
 * ForwardCPUAlgorithm; the citation on each declaration names the member it stands in for.
 *
 * Conventions: plain C types, no exceptions.  Every function returns RL_OK (0) or a negative
 * RL_E* status; rl_bh_last_error() returns a message for the last failure on the context.  A
 * context owns all its device memory and its HIP stream; calls on one context must not overlap.
 * Host arrays are caller owned and read (or written) only during the call.
 */
#ifndef RELEARN_BH_H
#define RELEARN_BH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RL_OK = 0,
    RL_EINVAL = -1,    /* bad argument: sizes, theta outside (0, 0.5], sigma <= 0, state order */
    RL_EHIP = -2,      /* a HIP runtime call failed */
    RL_EOOM = -3,      /* device allocation failed */
    RL_EOUTSIDE = -4,  /* a position lies outside the simulation box */
    RL_EDUPPOS = -5,   /* two neurons share a position, or are closer than 42 octree levels resolve */
    RL_ESTATE = -6,    /* call out of order (e.g. search before prepare) */
    RL_ELIMIT = -7,    /* more than 2^28 neurons or tree nodes */
};

enum {
    RL_PREC_FP64 = 0,        /* all arithmetic in IEEE binary64, same operation order as the CPU path */
    RL_PREC_FP32_MIXED = 1,  /* FP32 distances and Gaussian; FP64 centroids, sums and prefix sums */
};

enum {
    RL_FLAG_EXACT_ORDER = 1,      /* per-source search with sums in the CPU path's list order (slower) */
    RL_FLAG_WAVE_PER_SOURCE = 2,  /* FP32-mixed: one wavefront per source (the FP64 search) for every
                                     update.  By default an FP32-mixed update with at least 192 x 64 active
                                     sources per CU (3,145,728 on 256 CUs) takes the group walk (one lane
                                     per source, 64 neighbouring sources per wavefront), a smaller one a
                                     wavefront per source; the device chooses from its own count. */
    RL_FLAG_COUNT_WALK = 4,       /* FP32-mixed group walk: count its node steps and table slots into
                                     rl_bh_step_stats (a separate, slower instantiation of the walk) */
    RL_FLAG_WALK_ALWAYS = 8,      /* FP32-mixed: the group walk for every update, whatever its size */
};

enum {
    RL_SIGNAL_EXCITATORY = 0,
    RL_SIGNAL_INHIBITORY = 1,
};

typedef struct rl_bh_ctx rl_bh_ctx;

typedef struct {
    double box_min[3];   /* simulation box (bounding_box of BarnesHut's constructor) */
    double box_max[3];
    double theta;        /* acceptance criterion, 0 < theta <= 0.5 */
    double sigma;        /* Gaussian kernel sigma, > 0 */
    double mu;           /* Gaussian kernel mu */
    uint64_t seed;       /* seed of the counter-based draw stream */
    int32_t precision;   /* RL_PREC_* */
    int32_t device;      /* HIP device ordinal */
    int32_t flags;       /* RL_FLAG_* bits */
    int32_t reserved;
} rl_bh_config;

typedef struct {
    uint64_t searches;          /* find_target_neuron calls (vacant axons searched until the first failure) */
    uint64_t requests;          /* searches that found a target (synapse creation requests) */
    uint64_t created;           /* requests that were accepted by the commit */
    uint64_t node_tests;        /* acceptance tests executed on the device */
    uint64_t node_tests_ref;    /* acceptance tests the CPU path performs for the same searches */
    double ms_prepare;          /* leaf fill + upward pass */
    double ms_search;           /* target search */
    double ms_commit;           /* conflict resolution + vacancy commit */
    double ms_search_kernel;    /* the dominant search kernel alone (cooperative or per-lane) */
    double ms_total;            /* the whole update, or the whole replayed period (rl_bh_graph_launch) */
    uint64_t deleted;           /* synapses deleted by the period's element commit (rl_bh_graph_launch) */
    uint64_t net_capacity;      /* network graph capacity after the replay (rl_bh_graph_launch) */
    uint64_t net_regrown;       /* 1: the replay's synapses did not fit the network graph, which grew and
                                   took them; the period was captured again (rl_bh_graph_launch) */
    uint64_t lane_sources;      /* sources whose level-1 list went to the per-lane search: a cooperative
                                   slot table that overflowed, or a Gaussian sum of 0 (the FP64 re-evaluation
                                   of an FP32 sum of 0, or the n/r fallback of Kernel.h:109-126) */
    uint64_t slow_descents;     /* descents below level 1 handed to the per-lane traversal (expansion or a
                                   Gaussian sum of 0 below the picked node) */
    uint64_t walk_steps;        /* FP32-mixed group walk with RL_FLAG_COUNT_WALK: node steps (one node evaluated
                                   for the up to 64 sources of a wavefront); 0 without the flag */
    uint64_t walk_slots;        /* FP32-mixed group walk with RL_FLAG_COUNT_WALK: expanded nodes that got a slot
                                   of the per-source subtree sums; 0 without the flag */
    uint64_t walk_table_bytes;  /* FP32-mixed group walk with RL_FLAG_COUNT_WALK: bytes stored in the slot tables
                                   (8 per running sum of a source that expanded a node with an expanded
                                   child, 32 per child-entry row of such a node); 0 without the flag */
} rl_bh_step_stats;

/* Algorithm's constructor: BarnesHut(bounding_box, sfc, theta) (BarnesHut.h:54-60) with a
 * GaussianDistributionKernel(mu, sigma) (source/algorithm/Kernel/Gaussian.h:47-52). */
int rl_bh_create(const rl_bh_config* cfg, rl_bh_ctx** out);

/* Destroys the context and frees its device memory. */
int rl_bh_destroy(rl_bh_ctx* ctx);

/* Message of the last failing call on ctx (or of rl_bh_create when ctx is NULL). */
const char* rl_bh_last_error(const rl_bh_ctx* ctx);

/* Algorithm::init(number_neurons) (Algorithm.h:111) = OctreeAlgorithm::init
 * (source/algorithm/Internal/OctreeAlgorithm.h:58-67): builds the octree of the n neurons.
 * xyz: n x 3 AoS positions (NeuronsExtraInfo positions); signal: n signal types of the axons
 * (SynapticElements::get_signal_types), RL_SIGNAL_*. */
int rl_bh_init(rl_bh_ctx* ctx, uint64_t n, const double* xyz, const uint8_t* signal);

/* Ensemble of independent networks in one context, advanced together by every call below (batched
 * plasticity updates; no counterpart in the reference, which runs one network per process).
 * Network g owns the neurons [offsets[g], offsets[g+1]) (offsets[0] = 0, every network non-empty)
 * and the simulation box boxes[6g .. 6g+5] (min xyz, max xyz).  Each network gets its own octree
 * (one forest, built in one pass), its own fixed-point scale, and its own draw stream: seed
 * rl_bh_network_seed(cfg.seed, g) addressed by neuron ids local to the network.  Network g therefore
 * reproduces a one-network context created with that seed on its neurons.  rl_bh_init is the case
 * num_networks = 1 with the box of the configuration. */
int rl_bh_init_networks(rl_bh_ctx* ctx, uint32_t num_networks, const uint64_t* offsets, const double* boxes,
                        const double* xyz, const uint8_t* signal);

/* Algorithm::create_neurons(creation_count) (Algorithm.h:117) = OctreeAlgorithm::create_neurons
 * (source/algorithm/Internal/OctreeAlgorithm.h:73-85), run on a creation event (Neurons::create_neurons,
 * source/neurons/Neurons.cpp:324-343, from Simulation.cpp:425): appends `count` neurons with ids
 * n .. n + count - 1 at positions xyz (count x 3 AoS, inside the box) with signal types `signal`,
 * inserts them into the octree and re-initialises the leaves.  The octree of distinct positions does not
 * depend on the insertion order, so the tree equals the one rl_bh_init builds over all neurons.  The
 * new neurons are enabled, with no vacant or connected elements.  When the context holds neuron state
 * (rl_neurons_init), the new neurons start with x, u, ca, target_ca (count each; NULL x, u, ca give 0,
 * target_ca is then required) and grown (3 x count, type-major; NULL gives 0), as the reference's
 * calculators initialise them (CalciumCalculator::create_neurons, ElementBaseBase::create_neurons), and
 * the state of the existing neurons (deltas and fired status included) carries over; per-neuron element
 * parameters (rl_neurons_set_element_params) of the new neurons take the per-type values of
 * rl_step_params.  A network graph (rl_net_init) keeps its edges and capacity; the new neurons have
 * none (NetworkGraph::create_neurons).  A captured period is released.  One-network contexts only
 * (RL_ESTATE otherwise).  Positions and signal types are checked before anything changes; a failure
 * after that (RL_EDUPPOS, RL_EOOM) leaves the context to be initialised again. */
int rl_bh_create_neurons(rl_bh_ctx* ctx, uint64_t count, const double* xyz, const uint8_t* signal, const double* x,
                         const double* u, const double* ca, const double* target_ca, const double* grown);

/* Draw-stream seed of network g of an ensemble (network 0 keeps the configured seed). */
uint64_t rl_bh_network_seed(uint64_t seed, uint32_t network);

/* Algorithm::disable_neurons / NeuronsExtraInfo update flags: enabled[i] = 0 disables neuron i
 * (its leaf gets zero vacancies and it searches no targets).  NULL enables all. */
int rl_bh_set_enabled(rl_bh_ctx* ctx, const uint8_t* enabled);

/* Algorithm::prepare_update_connectivity(signal_types, vacant_axons, vacant_excitatory_dendrites,
 * vacant_inhibitory_dendrites) (Algorithm.h:136-140; OctreeAlgorithm::update_tree,
 * OctreeAlgorithm.h:100-112): uploads the vacancy spans, fills the leaves and runs the upward pass. */
int rl_bh_prepare_update_connectivity(rl_bh_ctx* ctx, const uint32_t* vacant_axons,
                                      const uint32_t* vacant_excitatory_dendrites,
                                      const uint32_t* vacant_inhibitory_dendrites);

/* Same as rl_bh_prepare_update_connectivity on the vacancies already resident on the device
 * (those of the last upload or commit). */
int rl_bh_prepare_resident(rl_bh_ctx* ctx);

/* rl_bh_prepare_update_connectivity for callers whose vacancy spans live in device memory of the
 * context's GPU (e.g. synaptic elements kept on the device): device-to-device copies, then leaf fill
 * and upward pass.  The pointers are read during the call only. */
int rl_bh_prepare_update_connectivity_device(rl_bh_ctx* ctx, const uint32_t* d_vacant_axons,
                                             const uint32_t* d_vacant_excitatory_dendrites,
                                             const uint32_t* d_vacant_inhibitory_dendrites);

/* BarnesHut::find_target_neurons (BarnesHut.cpp:36-94): one search per vacant axon of every enabled
 * neuron, stopping at a neuron's first failure.  step is the plasticity-update counter that
 * addresses the draw stream.  Returns the number of requests in *num_requests (may be NULL). */
int rl_bh_find_target_neurons(rl_bh_ctx* ctx, uint32_t step, uint64_t* num_requests);

/* Copies the requests of the last search, ordered by (source neuron, axon ordinal j), i.e. the
 * order a single-threaded CPU run appends them (SynapseCreationRequest{target, source, signal}). */
int rl_bh_get_requests(rl_bh_ctx* ctx, uint64_t capacity, uint32_t* target, uint32_t* source,
                       uint8_t* signal, uint32_t* axon_ordinal);

/* Per-slot diagnostics of the last search: slots are laid out by neuron id, vacant_axons[i] slots
 * per neuron starting at the exclusive prefix sum; node_tests[slot] counts the acceptance tests the
 * CPU path performs for that search, target[slot] the found target (0xFFFFFFFF when none). */
int rl_bh_get_search_slots(rl_bh_ctx* ctx, uint64_t capacity, uint32_t* target, uint32_t* node_tests,
                           uint32_t* count_per_neuron);

/* ForwardConnector::process_requests + process_responses (source/algorithm/Connector.cpp:33-197):
 * resolves conflicts in the order given by `order` (a permutation of the request indices, the
 * shuffled order of Connector.cpp:71) or, when order is NULL, by a counter-based priority
 * (seed, source, step, j); commits the vacancies on the device.  *created gets the number of synapses.
 * The requests of one search are committed once: a second call before the next search (or after a
 * replay with stats, which commits its own requests) returns RL_ESTATE and changes nothing; the
 * requests and accept flags stay readable. */
int rl_bh_commit(rl_bh_ctx* ctx, uint32_t step, const uint64_t* order, uint64_t* created);

/* Accept flags of the last commit (1 = SynapseCreationResponse::Succeeded), request order. */
int rl_bh_get_commit_result(rl_bh_ctx* ctx, uint64_t capacity, uint8_t* ok);

/* Algorithm::update_connectivity(number_neurons) (Algorithm.h:149) for the forward algorithm
 * (source/algorithm/Internal/ExchangingAlgorithm.h:58-87): prepare (resident vacancies) + search +
 * commit with counter priorities.  stats may be NULL. */
int rl_bh_update_connectivity(rl_bh_ctx* ctx, uint32_t step, rl_bh_step_stats* stats);

/* Copies the device-resident vacancies (axon, excitatory dendrite, inhibitory dendrite). */
int rl_bh_get_vacancies(rl_bh_ctx* ctx, uint32_t* vacant_axons, uint32_t* vacant_excitatory_dendrites,
                        uint32_t* vacant_inhibitory_dendrites);

/* Number of octree nodes and the depth (deepest level, root = 0) after rl_bh_init. */
int rl_bh_tree_info(rl_bh_ctx* ctx, uint64_t* num_nodes, uint32_t* depth);

/* Octree export in breadth-first order (level, then Morton order), the layout of the device tree.
 * Any pointer may be NULL.  pos arrays are 3 per node; has_* = 1 when the node has a position for
 * that dendrite type (leaves always, inner nodes when the count is non-zero). */
int rl_bh_export_tree(rl_bh_ctx* ctx, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant,
                      double* box_lo, double* box_hi, double* side, uint32_t* cnt_exc, uint32_t* cnt_inh,
                      uint8_t* has_exc, uint8_t* has_inh, double* pos_exc, double* pos_inh);

/* Milliseconds of the last search kernel(s), measured with HIP events on the context's stream. */
int rl_bh_last_search_ms(rl_bh_ctx* ctx, double* ms);

/* Diagnostic: the level-1 candidate list of each given source after rl_bh_prepare_* (the accepted
 * nodes of BarnesHutBase::get_nodes_to_consider at the root, BarnesHutBase.h:129-206, in list
 * order) with each candidate's probability as the context's precision computes it
 * (Kernel::calculate_attractiveness_to_connect, Kernel.h:61-75).  Source i's list is written to
 * nodes/probabilities[i * capacity ...]; counts[i] is its full length (entries beyond capacity are
 * not written). */
int rl_bh_candidate_lists(rl_bh_ctx* ctx, uint64_t num_sources, const uint32_t* sources, uint64_t capacity,
                          uint32_t* nodes, double* probabilities, uint32_t* counts);

/* Counters of the last search: searches, requests, node_tests, node_tests_ref, ms_search,
 * ms_search_kernel, lane_sources and slow_descents are filled (created and the other times are 0). */
int rl_bh_last_search_stats(rl_bh_ctx* ctx, rl_bh_step_stats* stats);

/* ------------------------------------------------------------------ neuron step */

typedef struct {
    double a, b, c, d;       /* Izhikevich parameters (models/izhikevich/Parameters.h:41) */
    double v_spike;
    double k1, k2, k3;
    uint32_t h;              /* integration substeps (NeuronModelBase.h:241) */
    uint32_t input_mode;     /* 0: per-step input array supplied by the caller, 1: counter-based Normal */
    double input_mean, input_stddev;   /* background input N(mean, stddev) for input_mode 1 */
    double tau_ca, beta_ca;  /* calcium decay and per-spike increment */
    double eta[3];           /* minimum calcium to grow: axon, excitatory, inhibitory dendrite */
    double nu[3];            /* growth rates */
    double synapse_conductance;  /* conductance per unit synapse weight (SynapticEquallyWeightedActivityInput) */
    uint32_t synaptic_input; /* 1: add the synaptic input of the network graph (rl_net_init) */
    uint32_t reserved;
    double retract_ratio[3]; /* vacant retract ratio per element type (--retract-ratio, relearn.cpp:617-618;
                                ElementBaseBase::update, ElementBaseBase.h:420-490); 0 in the reference's default */
} rl_step_params;

/* Allocates and uploads the per-neuron state of the step kernel (IzhikevichModel x/u, calcium,
 * target calcium, grown elements and deltas; 3 element types type-major).  NULL arrays start at 0
 * (target_ca must be given). */
int rl_neurons_init(rl_bh_ctx* ctx, const rl_step_params* p, const double* x, const double* u,
                    const double* ca, const double* target_ca, const double* grown,
                    const uint32_t* connected);

/* One simulation step of NeuronModel::update_electrical_activity + calcium + synaptic element
 * deltas (Simulation.cpp:443-453): Izhikevich (IzhikevichModelCPU.cpp:75-125), calcium
 * (CalciumCalculatorCPU.cpp:40-103), growth curve (SynapticElementsCPU.cpp:25-49).  input: n values
 * for input_mode 0, else NULL.  step addresses the counter-based input draws. */
int rl_neurons_step(rl_bh_ctx* ctx, uint32_t step, const double* input);

/* ElementBaseBase::update (ElementBaseBase.h:420-487) for the three element types: applies the
 * deltas to grown/vacant/connected counts at a plasticity step.  deleted (3*n, may be NULL) gets the
 * number of bound elements each neuron lost per type.  With a network graph (rl_net_init) the
 * synapses those elements held are deleted as SynapseDeletionFinderCPU::delete_synapses does
 * (SynapseDeletionFinderCPU.cpp): the types are committed in the order axons, excitatory, inhibitory
 * dendrites, each followed by drawing the lost number of synapse units uniformly without
 * replacement (RandomSynapseDeletionFinder, SynapseDeletionFinder.cpp:53-69; counter-based draw
 * addressed by plasticity update `update`), lowering the pairs' weights and disconnecting one element
 * of each partner (ElementBaseBase::disconnect_elements, ElementBaseBase.h:311-319).
 * rl_neurons_commit_elements uses update index 0. */
int rl_neurons_commit_elements(rl_bh_ctx* ctx, uint32_t* deleted);
int rl_neurons_commit_elements_at(rl_bh_ctx* ctx, uint32_t update, uint32_t* deleted);

/* Growth rates per network of an ensemble (nu[3g + t] for element type t of network g: axons,
 * excitatory, inhibitory dendrites), replacing rl_step_params.nu for those networks; NULL restores
 * rl_step_params.nu everywhere (SynapticElements growth rate, SynapticElementsBase.h:617-625). */
int rl_neurons_set_network_growth(rl_bh_ctx* ctx, const double* nu);

/* Per-neuron synaptic element parameters, as the reference's --synaptic-elements-file sets them
 * (SynapticElementsIO::load_function_from_file, source/io/SynapticElementsIO.h:60; applied in
 * relearn.cpp:1158-1170 through set_minimum_calcium_calculator and set_vacant_retract_ratio_calculator):
 * min_calcium[t * n + i] replaces rl_step_params.eta[t] for neuron i in the growth curve
 * (SynapticElementsCPU.cpp:25-49), retract_ratio[t * n + i] replaces rl_step_params.retract_ratio[t] in
 * the element commit (ElementBaseBase.h:420-490); t = axon, excitatory, inhibitory dendrite.  Either
 * pointer may be NULL, which restores the per-type values of rl_step_params for that parameter.  Values
 * must be finite; retract ratios lie in [0, 1]. */
int rl_neurons_set_element_params(rl_bh_ctx* ctx, const double* min_calcium, const double* retract_ratio);

/* Copies the neuron state: any pointer may be NULL; delta/grown are 3*n (type-major). */
int rl_neurons_get(rl_bh_ctx* ctx, double* x, double* u, double* ca, uint8_t* fired, double* delta,
                   double* grown, uint32_t* connected);

/* ------------------------------------------------------------------ network graph and synaptic input */

/* Network graph of the synaptic input (NetworkGraph's local in-edges, source/neurons/NetworkGraph.h):
 * per target neuron its in-edges (source, integer weight), CSR by target: edges
 * [in_offsets[t], in_offsets[t+1]) of the n + 1 offsets.  in_offsets NULL starts with no edge.  From
 * then on every commit appends the synapses it creates, as Neurons::create_synapses adds them to the
 * graph (Neurons.cpp:440-477): weight +1 for an excitatory source, -1 for an inhibitory one
 * (Connector.cpp:33-121).  Repeated (source, target) pairs are merged into one entry whose weight
 * is their sum, and pairs of weight 0 are dropped, as NetworkGraph keeps them.  capacity: entries
 * held without reallocation; the eager calls grow it, and so does rl_bh_graph_launch with stats
 * (see there).  rl_step_params.synaptic_input = 1 adds synapse_conductance times the summed weights
 * of the in-edges whose source fired in the previous step to the input of each step
 * (SynapticEquallyWeightedActivityInputCPU::update_local_input,
 * SynapticEquallyWeightedActivityInputCPU.cpp:20-50, combined with the background input as
 * CombinedActivityInputCPU.h does); the sum is formed in 32-bit integers, so a neuron's summed
 * |weights| must stay below 2^31 (checked here, RL_EINVAL). */
int rl_net_init(rl_bh_ctx* ctx, uint64_t capacity, const uint64_t* in_offsets, const uint32_t* sources,
                const int32_t* weights);

/* Copies the network graph (CSR by target; the order of edges inside a target is unspecified).
 * With all three arrays NULL only num_edges is returned. */
int rl_net_get_edges(rl_bh_ctx* ctx, uint64_t capacity, uint64_t* in_offsets, uint32_t* sources, int32_t* weights,
                     uint64_t* num_edges);

/* ------------------------------------------------------------------ launch tuning */

typedef struct {
    uint32_t search_waves_per_cu;  /* persistent waves of the search per compute unit, 1..24 (the group
                                      walk's in an FP32-mixed context) */
    uint32_t search_chunk;         /* sources (groups for the walk) a search wave takes from the work
                                      queue at a time, 1..64 (0: 1) */
    uint64_t walk_min_sources;     /* FP32-mixed context without RL_FLAG_WALK_ALWAYS or
                                      RL_FLAG_WAVE_PER_SOURCE: active sources from which an update takes
                                      the group walk instead of a wavefront per source (default 192 x 64
                                      per CU; UINT64_MAX: never).  rl_bh_set_tuning: 0 keeps the current
                                      value. */
    uint32_t reserved[4];
} rl_bh_tuning;

/* Launch tuning of the search (no counterpart in the reference, whose CPU loop has none).  Results
 * do not depend on it; a captured period is released when it changes. */
int rl_bh_set_tuning(rl_bh_ctx* ctx, const rl_bh_tuning* tuning);
int rl_bh_get_tuning(rl_bh_ctx* ctx, rl_bh_tuning* tuning);

/* Times the search of the prepared update `step` (after rl_bh_prepare_*) at candidate tunings (waves
 * per CU 8..24, then work-queue chunks 1..8 at the best wave count) and keeps the fastest.  In an
 * FP32-mixed context (without RL_FLAG_WALK_ALWAYS or RL_FLAG_WAVE_PER_SOURCE) it first times the
 * group walk and a wavefront per source on the update and moves walk_min_sources so that updates with
 * this many active sources take the faster one; the walk's waves and chunk are then tuned when the walk
 * is the faster, and kept otherwise.  Only the search runs; the vacancies do not change, and afterwards
 * the search result of `step` is readable as after rl_bh_find_target_neurons.  best may be NULL. */
int rl_bh_autotune(rl_bh_ctx* ctx, uint32_t step, rl_bh_tuning* best);

/* ------------------------------------------------------------------ captured plasticity periods */

/* Captures one plasticity period of the simulation loop (Simulation::simulate, Simulation.cpp:443-459)
 * into a HIP graph on the context's stream: `steps` neuron steps (rl_neurons_step with counter-based
 * input, so rl_neurons_init with input_mode 1), then, when steps > 0, the element commit
 * (rl_neurons_commit_elements), then the connectivity update of rl_bh_update_connectivity on the
 * resident vacancies (ForwardCPUAlgorithm::update_connectivity, ExchangingAlgorithm.h:58-87: leaf
 * fill, upward pass, search, commit with counter priorities).  steps = 0 captures the connectivity
 * update alone.  The slot buffers are sized for max_vacant_axons vacant axons per update; an update
 * with more searches nothing, changes nothing and is reported as RL_ELIMIT.  Replays allocate
 * nothing and synchronise with nothing.  rl_bh_init*, rl_neurons_init and
 * rl_neurons_set_network_growth release the captured period, as does growing the slot buffers. */
int rl_bh_graph_create(rl_bh_ctx* ctx, uint32_t steps, uint64_t max_vacant_axons);

/* Replays the captured period: neuron steps first_step .. first_step + steps - 1 and connectivity
 * update index `update` (the step argument of rl_bh_update_connectivity).  With stats, waits for the
 * period, fills searches, requests, created, deleted, node_tests, node_tests_ref, ms_total (the other
 * times are 0) and the network graph fields, and makes the requests and accept flags readable
 * (rl_bh_get_requests, rl_bh_get_commit_result).  When the created synapses do not fit the network
 * graph (rl_net_init's capacity), the replay appends none of them; with stats the call then grows the
 * graph (at least doubling it), appends them, captures the period again with the same arguments and
 * sets net_regrown, so the period completes as an eager one would.  Returns RL_ELIMIT when this
 * replay exceeded max_vacant_axons (the update changed nothing: rl_bh_update_connectivity with the
 * same index completes the period) or an earlier replay without stats exceeded a capacity.  Without
 * stats (NULL) the replay is asynchronous on the context's stream. */
int rl_bh_graph_launch(rl_bh_ctx* ctx, uint32_t first_step, uint32_t update, rl_bh_step_stats* stats);

/* Releases the captured period (rl_bh_destroy does too). */
int rl_bh_graph_destroy(rl_bh_ctx* ctx);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the conflict resolution and
 * vacancy commit of rl_bh_commit in counter order, context-free, on caller device buffers.  It calls
 * no hipMalloc/hipFree: its grouping scratch is a static device array of the library (about 30 MiB per
 * GPU, held until the process ends), which limits a call to n <= 262144 neurons and m <= 1048576
 * requests (RL_ELIMIT beyond).  Request i is (d_target[i], d_source[i], d_signal[i], d_axon_ordinal[i]),
 * ids 0..n-1 of one network whose draw-stream seed is network_seed (rl_bh_network_seed(seed, 0) for a
 * context created with seed); requests are processed in the order of the counter-based priority
 * (network_seed, source, step, axon ordinal), ties by index, as rl_bh_commit with order NULL.  Self
 * requests and requests naming an id >= n fail.  Accepted requests take one vacant dendrite of the
 * target (signal 0 excitatory, 1 inhibitory) and one vacant axon of the source; the caller's requests
 * of a source must not outnumber its vacant axons.  d_connected_* may be NULL (not counted).
 * d_ok[i] = 1 for an accepted request.  All work is ordered on hip_stream (a hipStream_t, NULL = the
 * default stream); with created != NULL the call waits for it and stores the number of accepted
 * requests, with created == NULL it returns without waiting (capturable).  Calls that overlap on one
 * GPU share the scratch array and race; rl_bh_last_error(NULL) describes a failure. */
int rl_bh_commit_requests_noalloc(uint64_t n, uint64_t m, const uint32_t* d_target, const uint32_t* d_source,
                                  const uint8_t* d_signal, const uint32_t* d_axon_ordinal, uint64_t network_seed,
                                  uint32_t step, uint32_t* d_vacant_axons, uint32_t* d_vacant_excitatory_dendrites,
                                  uint32_t* d_vacant_inhibitory_dendrites, uint32_t* d_connected_axons,
                                  uint32_t* d_connected_excitatory_dendrites,
                                  uint32_t* d_connected_inhibitory_dendrites, uint8_t* d_ok, void* hip_stream,
                                  uint64_t* created);

#ifdef __cplusplus
}
#endif

#endif /* RELEARN_BH_H */
