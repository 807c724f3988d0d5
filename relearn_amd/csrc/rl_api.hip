// C ABI of librelearn_bh.so: context management and the Algorithm-shaped entry points.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "rl_internal.h"

namespace rl {

int fail(rl_bh_ctx* c, int code, const std::string& msg) {
    if (c != nullptr) {
        c->st.code = code;
        c->st.msg = msg;
    }
    return code;
}

namespace {
thread_local char g_global_error[256] = "";
}  // namespace

void set_global_error(const char* msg) {
    std::strncpy(g_global_error, msg, sizeof(g_global_error) - 1);
    g_global_error[sizeof(g_global_error) - 1] = 0;
}

int hip_fail(rl_bh_ctx* c, hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory) ? RL_EOOM : RL_EHIP;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

void destroy_graph(rl_bh_ctx* c) {
    if (c->graph_exec != nullptr) {
        (void)hipGraphExecDestroy(c->graph_exec);
        c->graph_exec = nullptr;
    }
    c->graph_steps = 0;
}

int ensure_tmp(rl_bh_ctx* c, size_t bytes) {
    if (bytes <= c->tmp_bytes && c->tmp != nullptr) {
        return RL_OK;
    }
    if (c->tmp != nullptr) {
        (void)hipFree(c->tmp);
        c->tmp = nullptr;
    }
    const size_t want = std::max<size_t>(bytes, 1 << 20);
    hipError_t e = hipMalloc(&c->tmp, want);
    if (e != hipSuccess) {
        c->tmp = nullptr;
        c->tmp_bytes = 0;
        return fail(c, RL_EOOM, "scratch allocation failed");
    }
    c->tmp_bytes = want;
    return RL_OK;
}

namespace {


void free_all(rl_bh_ctx* c) {
    destroy_graph(c);
    free_tree(c);
    neurons_free(c);
    net_free(c);
    dfree(&c->pos);
    dfree(&c->qpos);
    dfree(&c->signal);
    dfree(&c->enabled);
    for (int t = 0; t < 3; t++) {
        dfree(&c->vac[t]);
        dfree(&c->conn[t]);
    }
    dfree(&c->leaf_of);
    dfree(&c->order);
    dfree(&c->slot_off);
    dfree(&c->slot_target);
    dfree(&c->slot_tests);
    dfree(&c->found);
    dfree(&c->searched);
    dfree(&c->req_target);
    dfree(&c->req_source);
    dfree(&c->req_signal);
    dfree(&c->req_j);
    dfree(&c->req_ok);
    dfree(&c->wide);
    dfree(&c->req_off);
    dfree(&c->counters);
    dfree(&c->d_steps);
    dfree(&c->scan_part);
    dfree(&c->c_gcnt);
    dfree(&c->c_goff);
    dfree(&c->c_prio);
    dfree(&c->c_order);
    dfree(&c->c_sprio);
    dfree(&c->c_sidx);
    dfree(&c->coop_flag);
    dfree(&c->coop_work);
    dfree(&c->coop_cursor);
    dfree(&c->coop_overflow);
    dfree(&c->redo);
    dfree(&c->coop_node);
    dfree(&c->walk_flag);
    dfree(&c->walk_off);
    dfree(&c->walk_list);
    dfree(&c->walk_seg_groups);
    dfree(&c->walk_seg_goff);
    dfree(&c->walk_seg_first);
    dfree(&c->walk_seg_count);
    dfree(&c->walk_cum);
    dfree(&c->walk_child);
    dfree(&c->slow_items);
    dfree(&c->coop_fexp);
    dfree(&c->coop_mass);
    dfree(&c->net_off);
    dfree(&c->net_box);
    dfree(&c->net_scale);
    dfree(&c->net_seed);
    dfree(&c->net_k2q);
    dfree(&c->net_invq);
    dfree(&c->net_of);
    c->num_networks = 0;
    c->h_net_off.clear();
    c->h_net_box.clear();
    c->h_net_scale.clear();
    c->coop_cap = 0;
    c->coop_blocks = 0;
    c->scan_part_cap = 0;
    if (c->tmp != nullptr) {
        (void)hipFree(c->tmp);
        c->tmp = nullptr;
        c->tmp_bytes = 0;
    }
    c->slots_cap = 0;
    c->slots = 0;
    c->nreq = 0;
}

int set_device(rl_bh_ctx* c) {
    RL_HIP(c, hipSetDevice(c->cfg.device));
    return RL_OK;
}

}  // namespace
}  // namespace rl

using namespace rl;

extern "C" {

int rl_bh_create(const rl_bh_config* cfg, rl_bh_ctx** out) {
    g_global_error[0] = 0;
    if (out == nullptr || cfg == nullptr) {
        set_global_error("rl_bh_create: null argument");
        return RL_EINVAL;
    }
    *out = nullptr;
    // BarnesHut::BarnesHut (BarnesHut.h:55-60) and test_acceptance_criterion's theta range
    // (BarnesHutBase.h:74-77); GaussianDistributionKernel (Gaussian.h:50) requires sigma > 0.
    if (!(cfg->theta > 0.0) || !(cfg->theta <= 0.5) || !(cfg->sigma > 0.0) || !std::isfinite(cfg->mu)) {
        set_global_error("rl_bh_create: theta must be in (0, 0.5] and sigma > 0");
        return RL_EINVAL;
    }
    if (cfg->precision != RL_PREC_FP64 && cfg->precision != RL_PREC_FP32_MIXED) {
        set_global_error("rl_bh_create: unknown precision");
        return RL_EINVAL;
    }
    if ((cfg->flags & ~(RL_FLAG_EXACT_ORDER | RL_FLAG_WAVE_PER_SOURCE | RL_FLAG_COUNT_WALK | RL_FLAG_WALK_ALWAYS)) !=
        0) {
        set_global_error("rl_bh_create: unknown flag bits");
        return RL_EINVAL;
    }
    const double big = std::numeric_limits<double>::max() / 4.0;
    for (int d = 0; d < 3; d++) {
        if (!(cfg->box_min[d] <= cfg->box_max[d]) || !(std::fabs(cfg->box_min[d]) < big) ||
            !(std::fabs(cfg->box_max[d]) < big)) {
            set_global_error("rl_bh_create: invalid simulation box");
            return RL_EINVAL;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_global_error("rl_bh_create: no HIP device");
        return RL_EHIP;
    }
    if (cfg->device < 0 || cfg->device >= ndev) {
        set_global_error("rl_bh_create: device ordinal out of range");
        return RL_EINVAL;
    }
    auto* c = new rl_bh_ctx();
    c->cfg = *cfg;
    if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
        set_global_error("rl_bh_create: stream/event creation failed");
        delete c;
        return RL_EHIP;
    }
    *out = c;
    return RL_OK;
}

int rl_bh_destroy(rl_bh_ctx* c) {
    if (c == nullptr) {
        return RL_OK;
    }
    (void)hipSetDevice(c->cfg.device);
    if (c->stream != nullptr) {
        (void)hipStreamSynchronize(c->stream);
    }
    free_all(c);
    if (c->ev0 != nullptr) {
        (void)hipEventDestroy(c->ev0);
    }
    if (c->ev2 != nullptr) {
        (void)hipEventDestroy(c->ev2);
    }
    if (c->ev3 != nullptr) {
        (void)hipEventDestroy(c->ev3);
    }
    if (c->ev1 != nullptr) {
        (void)hipEventDestroy(c->ev1);
    }
    if (c->stream != nullptr) {
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return RL_OK;
}

const char* rl_bh_last_error(const rl_bh_ctx* c) {
    if (c == nullptr) {
        return g_global_error;
    }
    return c->st.msg.c_str();
}

uint64_t rl_bh_network_seed(uint64_t seed, uint32_t network) {
    // network 0 keeps the context seed, so a one-network ensemble equals rl_bh_init
    return network == 0 ? seed : seed ^ (static_cast<uint64_t>(network) * 0x9E3779B97F4A7C15ULL);
}

namespace {

// Shared body of rl_bh_init and rl_bh_init_networks: validates, uploads and builds the forest.
int init_networks(rl_bh_ctx* c, uint32_t G, const uint64_t* offsets, const double* boxes, const double* xyz,
                  const uint8_t* signal) {
    const uint64_t n = offsets[G];
    if (n > 0 && (xyz == nullptr || signal == nullptr)) {
        return fail(c, RL_EINVAL, "rl_bh_init: null positions or signal types");
    }
    if (n > kLinkMask) {
        return fail(c, RL_ELIMIT, "rl_bh_init: more than 2^28 neurons");
    }
    if (offsets[0] != 0) {
        return fail(c, RL_EINVAL, "rl_bh_init_networks: offsets[0] must be 0");
    }
    std::vector<uint32_t> net_of(n);
    std::vector<double> scale(G);
    for (uint32_t g = 0; g < G; g++) {
        if (n > 0 && offsets[g + 1] <= offsets[g]) {
            return fail(c, RL_EINVAL, "rl_bh_init_networks: every network needs at least one neuron");
        }
        const double* b = boxes + 6 * static_cast<uint64_t>(g);
        double ext_max = 0.0;
        for (int d = 0; d < 3; d++) {
            if (!std::isfinite(b[d]) || !std::isfinite(b[3 + d]) || !(b[d] <= b[3 + d]) ||
                std::fabs(b[d]) > DBL_MAX / 4 || std::fabs(b[3 + d]) > DBL_MAX / 4) {
                return fail(c, RL_EINVAL, "rl_bh_init_networks: network box must be finite with min <= max");
            }
            ext_max = std::max(ext_max, b[3 + d] - b[d]);
        }
        // one fixed-point scale for all axes of the network (the largest extent spans 2^31 - 1
        // units), so squared distances and squared cell sides compare in the same units
        scale[g] = ext_max > 0.0 ? 2147483647.0 / ext_max : 0.0;
        for (uint64_t i = offsets[g]; i < offsets[g + 1]; i++) {
            net_of[i] = g;
            for (int d = 0; d < 3; d++) {
                const double v = xyz[3 * i + d];
                if (!(b[d] <= v && v <= b[3 + d])) {
                    return fail(c, RL_EOUTSIDE, "rl_bh_init: position outside the simulation box");
                }
            }
            if (signal[i] > 1) {
                return fail(c, RL_EINVAL, "rl_bh_init: signal type must be 0 or 1");
            }
        }
    }
    RL_TRY(set_device(c));
    free_all(c);
    c->n = n;
    c->prepared = false;
    c->searched_valid = false;
    c->num_networks = G;
    c->h_net_off.assign(offsets, offsets + G + 1);
    c->h_net_box.assign(boxes, boxes + 6 * static_cast<uint64_t>(G));
    c->h_net_scale = scale;
    std::vector<uint64_t> seeds(G);
    std::vector<float> k2q(G), invq(G);
    const double k2 = -1.4426950408889634 / (c->cfg.sigma * c->cfg.sigma);
    for (uint32_t g = 0; g < G; g++) {
        seeds[g] = rl_bh_network_seed(c->cfg.seed, g);
        const double iq = scale[g] > 0.0 ? 1.0 / scale[g] : 0.0;
        k2q[g] = static_cast<float>(k2 * iq * iq);
        invq[g] = static_cast<float>(iq);
    }
    RL_TRY(dalloc(c, &c->net_off, G + 1));
    RL_TRY(dalloc(c, &c->net_box, 6 * static_cast<size_t>(G)));
    RL_TRY(dalloc(c, &c->net_scale, G));
    RL_TRY(dalloc(c, &c->net_seed, G));
    RL_TRY(dalloc(c, &c->net_k2q, G));
    RL_TRY(dalloc(c, &c->net_invq, G));
    RL_TRY(dalloc(c, &c->net_of, n));
    RL_TRY(dalloc(c, &c->signal, n));
    RL_TRY(dalloc(c, &c->enabled, n));
    for (int t = 0; t < 3; t++) {
        RL_TRY(dalloc(c, &c->vac[t], n));
        RL_TRY(dalloc(c, &c->conn[t], n));
    }
    RL_TRY(dalloc(c, &c->slot_off, n + 1));
    RL_TRY(dalloc(c, &c->wide, n + 1));
    RL_TRY(dalloc(c, &c->req_off, n + 1));
    RL_TRY(dalloc(c, &c->counters, kCounters));
    RL_TRY(dalloc(c, &c->d_steps, 2));
    c->scan_part_cap = scan_scratch_entries(2 * n);
    RL_TRY(dalloc(c, &c->scan_part, c->scan_part_cap));
    RL_TRY(dalloc(c, &c->c_gcnt, 2 * n));
    RL_TRY(dalloc(c, &c->c_goff, 2 * n + 1));
    RL_TRY(dalloc(c, &c->found, n));
    RL_TRY(dalloc(c, &c->searched, n));
    RL_TRY(dalloc(c, &c->coop_flag, n));
    RL_TRY(dalloc(c, &c->coop_work, n));
    RL_TRY(dalloc(c, &c->coop_cursor, kCursorSlots));
    RL_TRY(dalloc(c, &c->coop_overflow, n));
    RL_TRY(dalloc(c, &c->redo, n));
    {
        int cus = 0;
        RL_HIP(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->cfg.device));
        c->cus = static_cast<uint32_t>(std::max(cus, 1));
        c->walk_min = walk_min_default(c);
        if (uses_walk(c)) {
            c->waves_per_cu = 24;   // the group walk's tables and LDS are sized for 24 waves per CU
        }
        c->coop_blocks = c->cus * c->waves_per_cu;
        c->coop_cap = 2048;
        if (!uses_walk(c) || walk_min_sources(c) > 0) {
            // the wave-per-source kernel (an FP32-mixed context with the group walk uses it for
            // updates with few active sources)
            const size_t slots = static_cast<size_t>(c->cus) * kMaxWavesPerCu * c->coop_cap;
            RL_TRY(dalloc(c, &c->coop_node, slots));
            RL_TRY(dalloc(c, &c->coop_fexp, slots));
            RL_TRY(dalloc(c, &c->coop_mass, slots));
        }
        if (uses_walk(c)) {
            RL_TRY(ensure_walk(c));
        }
    }
    hipStream_t s = c->stream;
    RL_HIP(c, hipMemcpyAsync(c->net_off, c->h_net_off.data(), (G + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemcpyAsync(c->net_box, c->h_net_box.data(), 6 * G * sizeof(double), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemcpyAsync(c->net_scale, scale.data(), G * sizeof(double), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemcpyAsync(c->net_seed, seeds.data(), G * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemcpyAsync(c->net_k2q, k2q.data(), G * sizeof(float), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemcpyAsync(c->net_invq, invq.data(), G * sizeof(float), hipMemcpyHostToDevice, s));
    RL_HIP(c, hipMemsetAsync(c->counters, 0, kCounters * sizeof(unsigned long long), s));
    RL_HIP(c, hipMemsetAsync(c->d_steps, 0, 2 * sizeof(uint32_t), s));
    if (n > 0) {
        RL_HIP(c, hipMemcpyAsync(c->net_of, net_of.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        RL_HIP(c, hipMemcpyAsync(c->signal, signal, n, hipMemcpyHostToDevice, s));
        RL_HIP(c, hipMemsetAsync(c->enabled, 1, n, s));
        for (int t = 0; t < 3; t++) {
            RL_HIP(c, hipMemsetAsync(c->vac[t], 0, n * sizeof(uint32_t), s));
            RL_HIP(c, hipMemsetAsync(c->conn[t], 0, n * sizeof(uint32_t), s));
        }
    }
    RL_HIP(c, hipStreamSynchronize(s));
    return build_tree(c, xyz);
}

}  // namespace

int rl_bh_init(rl_bh_ctx* c, uint64_t n, const double* xyz, const uint8_t* signal) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    const uint64_t off[2] = { 0, n };
    const double box[6] = { c->cfg.box_min[0], c->cfg.box_min[1], c->cfg.box_min[2],
                            c->cfg.box_max[0], c->cfg.box_max[1], c->cfg.box_max[2] };
    return init_networks(c, 1, off, box, xyz, signal);
}

int rl_bh_init_networks(rl_bh_ctx* c, uint32_t num_networks, const uint64_t* offsets, const double* boxes,
                        const double* xyz, const uint8_t* signal) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (num_networks == 0 || offsets == nullptr || boxes == nullptr) {
        return fail(c, RL_EINVAL, "rl_bh_init_networks: need at least one network, offsets and boxes");
    }
    if (num_networks > kLinkMask) {
        return fail(c, RL_ELIMIT, "rl_bh_init_networks: too many networks");
    }
    return init_networks(c, num_networks, offsets, boxes, xyz, signal);
}

int rl_bh_create_neurons(rl_bh_ctx* c, uint64_t count, const double* xyz, const uint8_t* signal, const double* x,
                         const double* u, const double* ca, const double* target_ca, const double* grown) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (c->n == 0 || c->num_networks != 1) {
        return fail(c, RL_ESTATE, "rl_bh_create_neurons: needs a context initialised by rl_bh_init");
    }
    if (count == 0 || xyz == nullptr || signal == nullptr) {
        return fail(c, RL_EINVAL, "rl_bh_create_neurons: count must be > 0 with positions and signal types");
    }
    const bool state = c->neurons.ready;
    if (state && target_ca == nullptr) {
        return fail(c, RL_EINVAL, "rl_bh_create_neurons: the neuron state needs the new neurons' target calcium");
    }
    RL_TRY(set_device(c));
    const uint64_t n0 = c->n;
    const uint64_t n1 = n0 + count;
    if (n1 > kLinkMask) {
        return fail(c, RL_ELIMIT, "rl_bh_create_neurons: more than 2^28 neurons");
    }
    hipStream_t s = c->stream;
    // Everything the context keeps per neuron is read back, extended and handed to the same
    // initialisation a fresh context runs: creation events are rare, and the octree of distinct
    // positions does not depend on the order of insertion (OctreeNode::insert), so rebuilding it over
    // all n0 + count neurons gives the tree that inserting the new ones would.
    std::vector<double> pos(3 * n0);
    std::vector<uint8_t> sig(n1), en(n1, 1);
    std::vector<uint32_t> vac(3 * n1, 0), conn(3 * n1, 0);
    RL_HIP(c, hipMemcpyAsync(pos.data(), c->pos, 3 * n0 * sizeof(double), hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(sig.data(), c->signal, n0, hipMemcpyDeviceToHost, s));
    RL_HIP(c, hipMemcpyAsync(en.data(), c->enabled, n0, hipMemcpyDeviceToHost, s));
    for (int t = 0; t < 3; t++) {
        RL_HIP(c, hipMemcpyAsync(vac.data() + t * n1, c->vac[t], n0 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipMemcpyAsync(conn.data() + t * n1, c->conn[t], n0 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    // neuron state: x, u, ca, target, fired status of the last step, grown and deltas (3 per neuron)
    const Neurons& ns = c->neurons;
    std::vector<double> sx, su, sca, stgt, sgrown, sdelta, smin, sret;
    std::vector<uint8_t> sfired;
    const rl_step_params params = ns.p;
    double* net_nu = nullptr;
    if (state) {
        sx.assign(n1, 0.0);
        su.assign(n1, 0.0);
        sca.assign(n1, 0.0);
        stgt.assign(n1, 0.0);
        sfired.assign(n1, 0);
        sgrown.assign(3 * n1, 0.0);
        sdelta.assign(3 * n1, 0.0);
        RL_HIP(c, hipMemcpyAsync(sx.data(), ns.x, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipMemcpyAsync(su.data(), ns.u, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipMemcpyAsync(sca.data(), ns.ca, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipMemcpyAsync(stgt.data(), ns.target_ca, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipMemcpyAsync(sfired.data(), ns.fired, n0, hipMemcpyDeviceToHost, s));
        for (int t = 0; t < 3; t++) {
            RL_HIP(c, hipMemcpyAsync(sgrown.data() + t * n1, ns.grown + t * n0, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
            RL_HIP(c, hipMemcpyAsync(sdelta.data() + t * n1, ns.delta + t * n0, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        if (ns.min_ca != nullptr || ns.retract != nullptr) {
            // new neurons take the per-type values of the step parameters
            smin.assign(3 * n1, 0.0);
            sret.assign(3 * n1, 0.0);
            for (int t = 0; t < 3; t++) {
                std::fill(smin.begin() + t * n1 + n0, smin.begin() + (t + 1) * n1, params.eta[t]);
                std::fill(sret.begin() + t * n1 + n0, sret.begin() + (t + 1) * n1, params.retract_ratio[t]);
                if (ns.min_ca != nullptr) {
                    RL_HIP(c, hipMemcpyAsync(smin.data() + t * n1, ns.min_ca + t * n0, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
                } else {
                    std::fill(smin.begin() + t * n1, smin.begin() + t * n1 + n0, params.eta[t]);
                }
                if (ns.retract != nullptr) {
                    RL_HIP(c, hipMemcpyAsync(sret.data() + t * n1, ns.retract + t * n0, n0 * sizeof(double), hipMemcpyDeviceToHost, s));
                } else {
                    std::fill(sret.begin() + t * n1, sret.begin() + t * n1 + n0, params.retract_ratio[t]);
                }
            }
        }
        net_nu = ns.net_nu;
    }
    double nu_net[3] = { 0.0, 0.0, 0.0 };
    if (net_nu != nullptr) {
        RL_HIP(c, hipMemcpyAsync(nu_net, net_nu, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    // network graph: the new neurons have no edges (NetworkGraph::create_neurons)
    const bool graph = c->net.ready;
    std::vector<uint64_t> off;
    std::vector<uint32_t> esrc;
    std::vector<int32_t> ew;
    uint64_t cap = 0;
    if (graph) {
        uint64_t edges = 0;
        RL_TRY(net_get(c, 0, nullptr, nullptr, nullptr, &edges));
        off.assign(n1 + 1, 0);
        esrc.assign(std::max<uint64_t>(edges, 1), 0);
        ew.assign(std::max<uint64_t>(edges, 1), 0);
        RL_TRY(net_get(c, edges, off.data(), esrc.data(), ew.data(), &edges));
        std::fill(off.begin() + n0 + 1, off.end(), edges);
        cap = c->net.cap;
    }
    RL_HIP(c, hipStreamSynchronize(s));
    std::vector<double> all(3 * n1);
    for (uint64_t i = 0; i < n0; i++) {
        for (int d = 0; d < 3; d++) {
            all[3 * i + d] = pos[d * n0 + i];
        }
    }
    std::memcpy(all.data() + 3 * n0, xyz, 3 * count * sizeof(double));
    std::memcpy(sig.data() + n0, signal, count);
    if (state) {
        for (uint64_t k = 0; k < count; k++) {
            sx[n0 + k] = x != nullptr ? x[k] : 0.0;
            su[n0 + k] = u != nullptr ? u[k] : 0.0;
            sca[n0 + k] = ca != nullptr ? ca[k] : 0.0;
            stgt[n0 + k] = target_ca[k];
            for (int t = 0; t < 3; t++) {
                sgrown[t * n1 + n0 + k] = grown != nullptr ? grown[t * count + k] : 0.0;
            }
        }
    }
    const uint64_t offs[2] = { 0, n1 };
    const double box[6] = { c->cfg.box_min[0], c->cfg.box_min[1], c->cfg.box_min[2],
                            c->cfg.box_max[0], c->cfg.box_max[1], c->cfg.box_max[2] };
    // validation of the new positions (box, signal types) happens here, before anything is released
    for (uint64_t k = 0; k < count; k++) {
        if (signal[k] > 1) {
            return fail(c, RL_EINVAL, "rl_bh_create_neurons: signal type must be 0 or 1");
        }
        for (int d = 0; d < 3; d++) {
            const double v = xyz[3 * k + d];
            if (!(box[d] <= v && v <= box[3 + d])) {
                return fail(c, RL_EOUTSIDE, "rl_bh_create_neurons: position outside the simulation box");
            }
        }
    }
    RL_TRY(init_networks(c, 1, offs, box, all.data(), sig.data()));
    RL_HIP(c, hipMemcpyAsync(c->enabled, en.data(), n1, hipMemcpyHostToDevice, s));
    for (int t = 0; t < 3; t++) {
        RL_HIP(c, hipMemcpyAsync(c->vac[t], vac.data() + t * n1, n1 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    if (state) {
        RL_TRY(neurons_init(c, &params, sx.data(), su.data(), sca.data(), stgt.data(), sgrown.data(), conn.data()));
        Neurons& nn = c->neurons;
        RL_HIP(c, hipMemcpyAsync(nn.delta, sdelta.data(), 3 * n1 * sizeof(double), hipMemcpyHostToDevice, s));
        RL_HIP(c, hipMemcpyAsync(nn.fired_buf[0], sfired.data(), n1, hipMemcpyHostToDevice, s));
        RL_HIP(c, hipStreamSynchronize(s));
        if (!smin.empty()) {
            RL_TRY(neurons_set_element_params(c, smin.data(), sret.data()));
        }
        if (net_nu != nullptr) {
            RL_TRY(neurons_set_network_growth(c, nu_net));
        }
    } else {
        for (int t = 0; t < 3; t++) {
            RL_HIP(c, hipMemcpyAsync(c->conn[t], conn.data() + t * n1, n1 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        }
        RL_HIP(c, hipStreamSynchronize(s));
    }
    if (graph) {
        RL_TRY(net_init(c, cap, off.data(), esrc.data(), ew.data()));
    }
    return RL_OK;
}

int rl_bh_set_enabled(rl_bh_ctx* c, const uint8_t* enabled) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    if (c->n == 0) {
        return RL_OK;
    }
    if (enabled == nullptr) {
        RL_HIP(c, hipMemsetAsync(c->enabled, 1, c->n, c->stream));
    } else {
        RL_HIP(c, hipMemcpyAsync(c->enabled, enabled, c->n, hipMemcpyHostToDevice, c->stream));
    }
    RL_HIP(c, hipStreamSynchronize(c->stream));
    c->prepared = false;
    return RL_OK;
}

int rl_bh_prepare_update_connectivity(rl_bh_ctx* c, const uint32_t* vacant_axons, const uint32_t* vacant_exc,
                                      const uint32_t* vacant_inh) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (c->n > 0 && (vacant_axons == nullptr || vacant_exc == nullptr || vacant_inh == nullptr)) {
        return fail(c, RL_EINVAL, "prepare_update_connectivity: null vacancy span");
    }
    RL_TRY(set_device(c));
    const uint32_t* h[3] = { vacant_axons, vacant_exc, vacant_inh };
    for (int t = 0; t < 3 && c->n > 0; t++) {
        RL_HIP(c, hipMemcpyAsync(c->vac[t], h[t], c->n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    }
    return rl_bh_prepare_resident(c);
}

int rl_bh_prepare_update_connectivity_device(rl_bh_ctx* c, const uint32_t* d_va, const uint32_t* d_ve,
                                             const uint32_t* d_vi) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (c->n > 0 && (d_va == nullptr || d_ve == nullptr || d_vi == nullptr)) {
        return fail(c, RL_EINVAL, "prepare_update_connectivity_device: null vacancy span");
    }
    RL_TRY(set_device(c));
    const uint32_t* d[3] = { d_va, d_ve, d_vi };
    for (int t = 0; t < 3 && c->n > 0; t++) {
        RL_HIP(c, hipMemcpyAsync(c->vac[t], d[t], c->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    }
    return rl_bh_prepare_resident(c);
}

int rl_bh_prepare_resident(rl_bh_ctx* c) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    RL_TRY(prepare_tree(c));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    c->prepared = true;
    c->searched_valid = false;
    return RL_OK;
}

int rl_bh_find_target_neurons(rl_bh_ctx* c, uint32_t step, uint64_t* num_requests) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    RL_TRY(find_targets(c, step));
    if (num_requests != nullptr) {
        *num_requests = c->nreq;
    }
    return RL_OK;
}

int rl_bh_get_requests(rl_bh_ctx* c, uint64_t capacity, uint32_t* target, uint32_t* source, uint8_t* signal,
                       uint32_t* axon_ordinal) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (!c->searched_valid) {
        return fail(c, RL_ESTATE, "get_requests: no search result");
    }
    if (capacity < c->nreq) {
        return fail(c, RL_EINVAL, "get_requests: capacity too small");
    }
    RL_TRY(set_device(c));
    const uint64_t m = c->nreq;
    hipStream_t s = c->stream;
    if (m > 0) {
        if (target) RL_HIP(c, hipMemcpyAsync(target, c->req_target, m * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (source) RL_HIP(c, hipMemcpyAsync(source, c->req_source, m * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (signal) RL_HIP(c, hipMemcpyAsync(signal, c->req_signal, m, hipMemcpyDeviceToHost, s));
        if (axon_ordinal) RL_HIP(c, hipMemcpyAsync(axon_ordinal, c->req_j, m * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    return RL_OK;
}

int rl_bh_get_search_slots(rl_bh_ctx* c, uint64_t capacity, uint32_t* target, uint32_t* node_tests,
                           uint32_t* count_per_neuron) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (!c->searched_valid) {
        return fail(c, RL_ESTATE, "get_search_slots: no search result");
    }
    if (capacity < c->slots) {
        return fail(c, RL_EINVAL, "get_search_slots: capacity too small");
    }
    RL_TRY(set_device(c));
    hipStream_t s = c->stream;
    if (c->slots > 0) {
        if (target) RL_HIP(c, hipMemcpyAsync(target, c->slot_target, c->slots * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (node_tests) RL_HIP(c, hipMemcpyAsync(node_tests, c->slot_tests, c->slots * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    if (count_per_neuron && c->n > 0) {
        RL_HIP(c, hipMemcpyAsync(count_per_neuron, c->found, c->n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    RL_HIP(c, hipStreamSynchronize(s));
    return RL_OK;
}

int rl_bh_commit(rl_bh_ctx* c, uint32_t step, const uint64_t* order, uint64_t* created) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    if (order != nullptr && c->nreq > 0) {
        std::vector<uint8_t> seen(c->nreq, 0);
        for (uint64_t k = 0; k < c->nreq; k++) {
            if (order[k] >= c->nreq || seen[order[k]] != 0) {
                return fail(c, RL_EINVAL, "commit: order is not a permutation of the request indices");
            }
            seen[order[k]] = 1;
        }
    }
    RL_TRY(commit(c, step, order, created));
    c->prepared = false;  // vacancies changed: the tree must be prepared again before the next search
    if (c->nreq > 0) {
        RL_TRY(net_append(c));
        RL_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RL_OK;
}

int rl_bh_get_commit_result(rl_bh_ctx* c, uint64_t capacity, uint8_t* ok) {
    if (c == nullptr || (ok == nullptr && c->nreq > 0)) {
        return RL_EINVAL;
    }
    if (capacity < c->nreq) {
        return fail(c, RL_EINVAL, "get_commit_result: capacity too small");
    }
    RL_TRY(set_device(c));
    if (c->nreq > 0) {
        RL_HIP(c, hipMemcpyAsync(ok, c->req_ok, c->nreq, hipMemcpyDeviceToHost, c->stream));
    }
    RL_HIP(c, hipStreamSynchronize(c->stream));
    return RL_OK;
}

int rl_bh_update_connectivity(rl_bh_ctx* c, uint32_t step, rl_bh_step_stats* stats) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    hipEvent_t e[4];
    for (auto& ev : e) {
        RL_HIP(c, hipEventCreate(&ev));
    }
    hipStream_t s = c->stream;
    int status = RL_OK;
    uint64_t created = 0;
    do {
        if (hipEventRecord(e[0], s) != hipSuccess) {
            status = fail(c, RL_EHIP, "event record");
            break;
        }
        if ((status = rl_bh_prepare_resident(c)) != RL_OK) break;
        if (hipEventRecord(e[1], s) != hipSuccess) {
            status = fail(c, RL_EHIP, "event record");
            break;
        }
        if ((status = find_targets(c, step)) != RL_OK) break;
        if (hipEventRecord(e[2], s) != hipSuccess) {
            status = fail(c, RL_EHIP, "event record");
            break;
        }
        if ((status = commit(c, step, nullptr, &created)) != RL_OK) break;
        c->prepared = false;
        if (c->nreq > 0 && (status = net_append(c)) != RL_OK) break;
        if (hipEventRecord(e[3], s) != hipSuccess || hipEventSynchronize(e[3]) != hipSuccess) {
            status = fail(c, RL_EHIP, "event record");
            break;
        }
        if (stats != nullptr) {
            float t01 = 0, t12 = 0, t23 = 0;
            (void)hipEventElapsedTime(&t01, e[0], e[1]);
            (void)hipEventElapsedTime(&t12, e[1], e[2]);
            (void)hipEventElapsedTime(&t23, e[2], e[3]);
            stats->searches = c->last_searches;
            stats->requests = c->nreq;
            stats->created = created;
            stats->node_tests = c->last_tests;
            stats->node_tests_ref = c->last_tests_ref;
            stats->ms_prepare = t01;
            stats->ms_search = t12;
            stats->ms_commit = t23;
            stats->ms_search_kernel = c->last_kernel_ms;
            stats->ms_total = t01 + t12 + t23;
            stats->lane_sources = c->last_lane_sources;
            stats->slow_descents = c->last_slow_descents;
            stats->walk_steps = c->last_walk_steps;
            stats->walk_slots = c->last_walk_slots;
            stats->walk_table_bytes = c->last_walk_bytes;
        }
    } while (false);
    for (auto& ev : e) {
        (void)hipEventDestroy(ev);
    }
    return status;
}

int rl_bh_get_vacancies(rl_bh_ctx* c, uint32_t* va, uint32_t* ve, uint32_t* vi) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    uint32_t* h[3] = { va, ve, vi };
    for (int t = 0; t < 3 && c->n > 0; t++) {
        if (h[t] != nullptr) {
            RL_HIP(c, hipMemcpyAsync(h[t], c->vac[t], c->n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        }
    }
    RL_HIP(c, hipStreamSynchronize(c->stream));
    return RL_OK;
}

int rl_bh_tree_info(rl_bh_ctx* c, uint64_t* num_nodes, uint32_t* depth) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (num_nodes) *num_nodes = c->tree.m;
    if (depth) *depth = c->tree.depth;
    return RL_OK;
}

int rl_bh_export_tree(rl_bh_ctx* c, uint8_t* level, uint8_t* nchild, uint32_t* link, uint8_t* octant, double* box_lo,
                      double* box_hi, double* side, uint32_t* cnt_exc, uint32_t* cnt_inh, uint8_t* has_exc,
                      uint8_t* has_inh, double* pos_exc, double* pos_inh) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return export_tree(c, level, nchild, link, octant, box_lo, box_hi, side, cnt_exc, cnt_inh, has_exc, has_inh,
                       pos_exc, pos_inh);
}

int rl_bh_last_search_ms(rl_bh_ctx* c, double* ms) {
    if (c == nullptr || ms == nullptr) {
        return RL_EINVAL;
    }
    *ms = c->last_search_ms;
    return RL_OK;
}

int rl_bh_last_search_stats(rl_bh_ctx* c, rl_bh_step_stats* st) {
    if (c == nullptr || st == nullptr) {
        return RL_EINVAL;
    }
    if (!c->searched_valid) {
        return fail(c, RL_ESTATE, "last_search_stats: no search result");
    }
    *st = rl_bh_step_stats{};
    st->searches = c->last_searches;
    st->requests = c->nreq;
    st->node_tests = c->last_tests;
    st->node_tests_ref = c->last_tests_ref;
    st->ms_search = c->last_search_ms;
    st->ms_search_kernel = c->last_kernel_ms;
    st->lane_sources = c->last_lane_sources;
    st->slow_descents = c->last_slow_descents;
    st->walk_steps = c->last_walk_steps;
    st->walk_slots = c->last_walk_slots;
    st->walk_table_bytes = c->last_walk_bytes;
    return RL_OK;
}

int rl_neurons_init(rl_bh_ctx* c, const rl_step_params* p, const double* x, const double* u, const double* ca,
                    const double* target_ca, const double* grown, const uint32_t* connected) {
    if (c == nullptr || p == nullptr || (c->n > 0 && target_ca == nullptr) || p->h == 0 || p->input_mode > 1) {
        return c == nullptr ? RL_EINVAL : fail(c, RL_EINVAL, "rl_neurons_init: bad arguments");
    }
    for (int t = 0; t < 3; t++) {
        if (!(p->retract_ratio[t] >= 0.0 && p->retract_ratio[t] <= 1.0)) {
            return fail(c, RL_EINVAL, "rl_neurons_init: retract ratios must lie in [0, 1]");
        }
    }
    RL_TRY(set_device(c));
    return neurons_init(c, p, x, u, ca, target_ca, grown, connected);
}

int rl_neurons_set_network_growth(rl_bh_ctx* c, const double* nu) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return neurons_set_network_growth(c, nu);
}

int rl_neurons_set_element_params(rl_bh_ctx* c, const double* min_calcium, const double* retract_ratio) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return neurons_set_element_params(c, min_calcium, retract_ratio);
}

int rl_neurons_step(rl_bh_ctx* c, uint32_t step, const double* input) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return neurons_step(c, step, input);
}

int rl_neurons_commit_elements(rl_bh_ctx* c, uint32_t* deleted) {
    return rl_neurons_commit_elements_at(c, 0, deleted);
}

int rl_neurons_commit_elements_at(rl_bh_ctx* c, uint32_t update, uint32_t* deleted) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    c->prepared = false;
    return neurons_commit(c, update, deleted);
}

int rl_neurons_get(rl_bh_ctx* c, double* x, double* u, double* ca, uint8_t* fired, double* delta, double* grown,
                   uint32_t* connected) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return neurons_get(c, x, u, ca, fired, delta, grown, connected);
}

int rl_bh_candidate_lists(rl_bh_ctx* c, uint64_t num_sources, const uint32_t* sources, uint64_t capacity,
                          uint32_t* nodes, double* probabilities, uint32_t* counts) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (num_sources > 0 && (sources == nullptr || counts == nullptr ||
                            (capacity > 0 && (nodes == nullptr || probabilities == nullptr)))) {
        return fail(c, RL_EINVAL, "rl_bh_candidate_lists: null array");
    }
    RL_TRY(set_device(c));
    return candidate_lists(c, num_sources, sources, capacity, nodes, probabilities, counts);
}

int rl_bh_set_tuning(rl_bh_ctx* c, const rl_bh_tuning* t) {
    if (c == nullptr || t == nullptr) {
        return RL_EINVAL;
    }
    if (t->search_waves_per_cu < 1 || t->search_waves_per_cu > kMaxWavesPerCu) {
        return fail(c, RL_EINVAL, "rl_bh_set_tuning: search_waves_per_cu must be in [1, 24]");
    }
    const uint32_t chunk = t->search_chunk == 0 ? 1U : t->search_chunk;
    if (chunk > 64) {
        return fail(c, RL_EINVAL, "rl_bh_set_tuning: search_chunk must be in [0, 64]");
    }
    const uint64_t walk_min = t->walk_min_sources == 0 ? c->walk_min : t->walk_min_sources;
    if (t->search_waves_per_cu != c->waves_per_cu || chunk != c->search_chunk || walk_min != c->walk_min) {
        destroy_graph(c);  // the captured launch has the old grid, chunk and kernel choice
    }
    c->search_chunk = chunk;
    c->waves_per_cu = t->search_waves_per_cu;
    c->coop_blocks = c->cus * c->waves_per_cu;
    c->walk_min = walk_min;
    return RL_OK;
}

int rl_bh_get_tuning(rl_bh_ctx* c, rl_bh_tuning* t) {
    if (c == nullptr || t == nullptr) {
        return RL_EINVAL;
    }
    *t = rl_bh_tuning{};
    t->search_waves_per_cu = c->waves_per_cu;
    t->search_chunk = c->search_chunk;
    t->walk_min_sources = c->walk_min;
    return RL_OK;
}

int rl_bh_autotune(rl_bh_ctx* c, uint32_t step, rl_bh_tuning* best) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (!c->prepared) {
        return fail(c, RL_ESTATE, "rl_bh_autotune: prepare the update first");
    }
    if (c->cus == 0) {
        return fail(c, RL_ESTATE, "rl_bh_autotune: no neurons (rl_bh_init first)");
    }
    RL_TRY(set_device(c));
    RL_TRY(find_targets(c, step));  // warm-up at the current setting
    if (c->last_active == 0) {
        // nothing searches in this update: nothing to time, the setting stays
        if (best != nullptr) {
            RL_TRY(rl_bh_get_tuning(c, best));
        }
        return RL_OK;
    }
    if (uses_walk(c) && (c->cfg.flags & RL_FLAG_WALK_ALWAYS) == 0) {
        // traversal width: the group walk (64 sources per wavefront) against a wavefront per source on
        // this update, each forced, twice, the faster run counting; the threshold moves so that updates
        // of this size take the faster kernel
        const uint64_t active = c->last_active;
        const uint64_t before = c->walk_min;
        double t_walk = 0.0, t_wave = 0.0;
        for (int k = 0; k < 2; k++) {
            c->walk_min = k == 0 ? 1ULL : UINT64_MAX;
            double ms = 0.0;
            for (int rep = 0; rep < 2; rep++) {
                RL_TRY(find_targets(c, step));
                ms = rep == 0 ? c->last_search_ms : std::min(ms, c->last_search_ms);
            }
            (k == 0 ? t_walk : t_wave) = ms;
        }
        const bool walk = t_walk < t_wave;
        c->walk_min = walk ? std::min(before, active) : std::max(before, active + 1);
        if (c->walk_min != before) {
            destroy_graph(c);
        }
        if (!walk) {
            // the wave-per-source kernel runs at its own fixed setting; the walk's setting stays
            RL_TRY(find_targets(c, step));
            if (best != nullptr) {
                RL_TRY(rl_bh_get_tuning(c, best));
            }
            return RL_OK;
        }
    }
    // the result is one of the candidates below (the fastest; the first of equals)
    rl_bh_tuning bt{};
    double best_ms = std::numeric_limits<double>::infinity();
    // each setting is timed twice and its faster run counts: single runs vary by 1-2%, as much as
    // neighbouring settings differ
    auto trial = [&](uint32_t waves, uint32_t chunk) -> int {
        rl_bh_tuning t{};
        t.search_waves_per_cu = waves;
        t.search_chunk = chunk;
        RL_TRY(rl_bh_set_tuning(c, &t));
        double ms = 0.0;
        for (int rep = 0; rep < 2; rep++) {
            RL_TRY(find_targets(c, step));
            ms = rep == 0 ? c->last_search_ms : std::min(ms, c->last_search_ms);
        }
        if (ms < best_ms) {
            best_ms = ms;
            bt = t;
        }
        return RL_OK;
    };
    // waves per CU at one source per grab, then the grab size at the best wave count
    for (uint32_t w : { 8U, 12U, 16U, 20U, 24U }) {
        RL_TRY(trial(w, 1U));
    }
    const uint32_t waves = bt.search_waves_per_cu;
    for (uint32_t k : { 2U, 4U, 8U }) {
        RL_TRY(trial(waves, k));
    }
    rl_bh_tuning t = bt;
    t.walk_min_sources = 0;   // keeps the width decision above
    RL_TRY(rl_bh_set_tuning(c, &t));
    RL_TRY(find_targets(c, step));  // leaves the search result of `step` readable, as find_target_neurons does
    if (best != nullptr) {
        RL_TRY(rl_bh_get_tuning(c, best));
    }
    return RL_OK;
}

int rl_bh_graph_create(rl_bh_ctx* c, uint32_t steps, uint64_t max_vacant_axons) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    constexpr uint32_t kMaxCapturedSteps = 100000;
    if (steps > kMaxCapturedSteps) {
        return fail(c, RL_EINVAL, "rl_bh_graph_create: more than 100000 steps in one period");
    }
    if (c->n == 0 || c->tree.m == 0) {
        return fail(c, RL_ESTATE, "rl_bh_graph_create: no neurons (rl_bh_init first)");
    }
    if (steps > 0 && (!c->neurons.ready || c->neurons.p.input_mode == 0)) {
        return fail(c, RL_ESTATE, "rl_bh_graph_create: neuron steps need rl_neurons_init with input_mode 1");
    }
    RL_TRY(set_device(c));
    destroy_graph(c);
    RL_TRY(ensure_slots(c, std::max<uint64_t>(max_vacant_axons, 1)));
    RL_TRY(normalize_fired(c));
    hipStream_t s = c->stream;
    RL_HIP(c, hipStreamSynchronize(s));
    RL_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int status = RL_OK;
    if (steps > 0) {
        status = enqueue_neuron_steps(c, steps);
        if (status == RL_OK) status = enqueue_elements_commit(c);
    }
    if (status == RL_OK) status = prepare_tree(c);
    if (status == RL_OK) status = enqueue_search(c);
    if (status == RL_OK) status = enqueue_commit(c, nullptr);
    if (status == RL_OK) status = enqueue_net_append(c);
    // the next period's steps push over the pairs as this period left them
    if (status == RL_OK && c->neurons.ready && c->neurons.p.synaptic_input != 0) status = enqueue_syn_sync(c, 0);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (status != RL_OK || e != hipSuccess) {
        if (g != nullptr) {
            (void)hipGraphDestroy(g);
        }
        return status != RL_OK ? status : hip_fail(c, e, "hipStreamEndCapture");
    }
    const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->graph_exec = nullptr;
        return hip_fail(c, ei, "hipGraphInstantiate");
    }
    c->graph_steps = steps;
    c->graph_max_vacant = max_vacant_axons;
    c->graph_appends = c->net.ready;
    c->prepared = false;
    c->searched_valid = false;
    return RL_OK;
}

int rl_bh_graph_launch(rl_bh_ctx* c, uint32_t first_step, uint32_t update, rl_bh_step_stats* stats) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    if (c->graph_exec == nullptr) {
        return fail(c, RL_ESTATE, "rl_bh_graph_launch: no captured period (rl_bh_graph_create)");
    }
    RL_TRY(set_device(c));
    hipStream_t s = c->stream;
    RL_TRY(normalize_fired(c));
    const bool synaptic = c->neurons.ready && c->neurons.p.synaptic_input != 0 && c->net.ready;
    if (synaptic) {
        RL_TRY(syn_prepare(c));
    }
    if (stats == nullptr && c->graph_appends) {
        // a replay without stats cannot grow the network graph after an append that does not fit (the
        // period would keep its committed vacancies without the synapses): it runs only while the graph
        // holds the most synapses the period can create (one per vacant axon of its capacity)
        const uint64_t bound = c->net.entries_bound + c->graph_max_vacant;
        if (bound > c->net.cap) {
            return fail(c, RL_ELIMIT, "rl_bh_graph_launch: without stats the network graph might not hold this "
                                      "period's synapses; launch with stats, which grows it");
        }
        c->net.entries_bound = bound;
    }
    RL_TRY(set_steps(c, first_step, update));
    if (stats != nullptr) {
        RL_HIP(c, hipEventRecord(c->ev0, s));
    }
    RL_HIP(c, hipGraphLaunch(c->graph_exec, s));
    if (synaptic) {
        c->net.syn_valid = true;   // the period ends by rebuilding the pushes over its pairs
    }
    c->prepared = false;
    c->searched_valid = false;
    c->nreq = 0;
    if (stats == nullptr) {
        return RL_OK;
    }
    RL_HIP(c, hipEventRecord(c->ev1, s));
    unsigned long long h[kCounters];
    RL_TRY(read_counters(c, h));
    if (c->graph_appends && c->n > 0) {
        uint64_t entries = 0;   // the graph's entries now (the stream is idle after the counters)
        RL_HIP(c, hipMemcpyAsync(&entries, c->net.off + c->n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        RL_HIP(c, hipStreamSynchronize(s));
        c->net.entries_bound = entries;
    }
    float ms = 0.0f;
    RL_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *stats = rl_bh_step_stats{};
    stats->searches = h[kCtrSearches];
    stats->requests = h[kCtrRequests];
    stats->created = h[kCtrCreated];
    stats->node_tests = h[kCtrTestsExec];
    stats->node_tests_ref = h[kCtrTestsRef];
    stats->deleted = h[kCtrDeleted];
    stats->ms_total = ms;
    c->slots = h[kCtrSlots];
    c->nreq = h[kCtrRequests];
    c->last_searches = h[kCtrSearches];
    c->last_tests = h[kCtrTestsExec];
    c->last_tests_ref = h[kCtrTestsRef];
    c->last_slow_descents = h[kCtrSlow];
    c->last_lane_sources = (c->cfg.flags & RL_FLAG_EXACT_ORDER) != 0 || c->tree.depth == 0 ? h[kCtrRedo] : h[kCtrOverflow];
    c->last_walk_steps = h[kCtrWalkSteps];
    c->last_walk_slots = h[kCtrWalkSlots];
    c->last_walk_bytes = h[kCtrWalkBytes];
    stats->lane_sources = c->last_lane_sources;
    stats->slow_descents = c->last_slow_descents;
    stats->walk_steps = c->last_walk_steps;
    stats->walk_slots = c->last_walk_slots;
    stats->walk_table_bytes = c->last_walk_bytes;
    c->searched_valid = true;
    c->committed = h[kCtrOverCap] == 0;   // the replay committed its requests (none when over capacity)
    if (h[kCtrOverCap] != 0) {
        return fail(c, RL_ELIMIT, "rl_bh_graph_launch: more vacant axons than max_vacant_axons; the update changed nothing");
    }
    if (h[kCtrNetOver] != 0) {
        // the period is complete except for the append of its synapses, whose counts are kept: grow
        // the network graph, append them, and capture the period again over the new arrays
        const uint32_t steps = c->graph_steps;
        const uint64_t max_vacant = c->graph_max_vacant;
        RL_TRY(net_append_retry(c));
        RL_TRY(rl_bh_graph_create(c, steps, max_vacant));
        stats->net_regrown = 1;
        c->slots = h[kCtrSlots];
        c->nreq = h[kCtrRequests];
        c->searched_valid = true;
        c->committed = true;
    }
    stats->net_capacity = c->net.ready ? c->net.cap : 0;
    if (h[kCtrSticky] > (h[kCtrNetOver] != 0 ? 1ULL : 0ULL)) {
        return fail(c, RL_ELIMIT, "rl_bh_graph_launch: an earlier replay without stats exceeded a capacity");
    }
    if (h[kCtrDelShort] != 0) {
        return fail(c, RL_EINVAL, "rl_bh_graph_launch: the network graph held fewer synapses than the element "
                                  "commit unbound (connected counts and graph disagree)");
    }
    return RL_OK;
}

int rl_net_init(rl_bh_ctx* c, uint64_t capacity, const uint64_t* in_offsets, const uint32_t* sources,
                const int32_t* weights) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return net_init(c, capacity, in_offsets, sources, weights);
}

int rl_net_get_edges(rl_bh_ctx* c, uint64_t capacity, uint64_t* in_offsets, uint32_t* sources, int32_t* weights,
                     uint64_t* num_edges) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    return net_get(c, capacity, in_offsets, sources, weights, num_edges);
}

int rl_bh_graph_destroy(rl_bh_ctx* c) {
    if (c == nullptr) {
        return RL_EINVAL;
    }
    RL_TRY(set_device(c));
    destroy_graph(c);
    return RL_OK;
}

}  // extern "C"
