// Fused per-step neuron update: background input, Izhikevich membrane update, calcium integrator and
// the growth-curve deltas of the three synaptic element types, one lane per neuron, in FP64 (the
// counter-based background draw, normal_input below, is FP32) and in the CPU path's operation order
// (files compiled with -ffp-contract=off):
//   IzhikevichModelCPU::update_activity   (source/neurons/models/izhikevich/IzhikevichModelCPU.cpp:75-125)
//   CalciumCalculatorCPU::update_current_calcium (source/neurons/calcium/CalciumCalculatorCPU.cpp:40-103)
//   update_number_elements_openmp + gaussian_growth_curve
//                                       (source/neurons/synaptic_elements/SynapticElementsCPU.cpp:25-49,
//                                        GrowthCurves.h:30-54)
// and, at plasticity steps, ElementBaseBase::update (ElementBaseBase.h:420-487) per element type.
#include <algorithm>
#include <cmath>
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

__device__ __forceinline__ double growth_curve(double current, double eta, double epsilon) {
#pragma clang fp contract(off)
    if (eta == epsilon) {
        return current == eta ? 0.0 : -1.0;
    }
    constexpr double factor = 1.6651092223153955127063292897904020952611777045288814583336582344;
    constexpr double factor_inv = 1.0 / factor;
    const double xi = (eta + epsilon) * 0.5;
    const double zeta = (eta - epsilon) * factor_inv;
    const double difference = current - xi;
    const double quotient = difference / zeta;
    const double product = quotient * quotient;
    return (2.0 * exp(-product)) - 1.0;
}

// Normal(mean, stddev) from two Philox words (tag 3) by the Box-Muller transform in FP32 on 24-bit
// uniforms: background noise needs no more (|z| <= 5.77, the mass beyond is 8e-9), and the FP32
// transcendentals take a third of what the FP64 ones cost in this kernel.  This input mode has no
// counterpart in the reference's stream (its normal input comes from Boost); the caller-array mode
// is the one to compare with the CPU path.
__device__ __forceinline__ double normal_input(uint64_t seed, uint32_t neuron, uint32_t step, double mean, double sd) {
    const uint4 o = philox4x32_10(make_uint4(neuron, step, 0U, 3U),
                                  make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U)));
    const float u1 = static_cast<float>((o.x >> 8U) + 1U) * 5.9604644775390625e-8f;  // (0, 1]
    const float u2 = static_cast<float>(o.y >> 8U) * 5.9604644775390625e-8f;         // [0, 1)
    const float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    return mean + sd * static_cast<double>(z);
}

struct StepArgs {
    uint64_t n;
    rl_step_params p;
    const uint32_t* net_of;      // network of each neuron, or null: one network
    const uint64_t* net_off;     // first neuron of each network
    const uint64_t* net_seed;    // draw seed of each network
    const double* net_nu;        // growth rates per network (3 each) or null: p.nu for all
    const double* min_ca;        // 3n per-neuron minimum calcium (type-major) or null: p.eta
    const uint32_t* step_base;   // device step base of a captured period (step = *step_base + step) or null
    uint32_t step;
    double* x;
    double* u;
    double* ca;
    const double* input;
    const uint8_t* fired_prev;   // fired status of the previous step (disabled carry)
    uint8_t* fired;
    const double* target_ca;     // per-neuron target calcium, or null: target_value for all
    double target_value;
    double* delta;
    const uint8_t* enabled;
    // synaptic input, or null: acc_cur holds the summed weights that the sources which fired in the
    // previous step pushed; a neuron that fires in this step pushes its out-pairs into acc_next
    int32_t* acc_cur;
    int32_t* acc_next;
    const uint64_t* out_off;
    const uint32_t* out_tgt;
    const int32_t* out_w;
    double conductance;
    // 1 / h and -1 / tau_ca, divided once on the host (IEEE division is correctly rounded on both
    // sides, so the values are the ones the reference's loops use)
    double scale;
    double tau_inv;
};

// A neuron that fired adds the weight of each of its out-pairs to the target's input of the next
// step (integer atomics: the sums do not depend on their order).  Few neurons fire per step, so
// this replaces a gather over every in-edge of every neuron.
// Pairs are read four at a time, so that their loads overlap.
__device__ __forceinline__ void push_out_pairs(const StepArgs& a, uint64_t i) {
    const uint64_t e0 = a.out_off[i];
    const uint64_t e1 = a.out_off[i + 1];
    for (uint64_t e = e0; e < e1; e += 4) {
        uint32_t t[4];
        int32_t w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool in = e + k < e1;
            t[k] = in ? a.out_tgt[e + k] : 0U;
            w[k] = in ? a.out_w[e + k] : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (w[k] != 0) {
                atomicAdd(&a.acc_next[t[k]], w[k]);
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_step(StepArgs a) {
#pragma clang fp contract(off)
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= a.n) {
        return;
    }
    const uint64_t n = a.n;
    const rl_step_params& p = a.p;
    // every per-neuron load first, so that they are in flight together instead of one round trip
    // after another
    const uint8_t en = a.enabled[i];
    const uint8_t fprev = a.fired_prev[i];
    const uint32_t g = a.net_of != nullptr ? a.net_of[i] : 0U;
    double xv = a.x[i];
    double uv = a.u[i];
    double c = a.ca[i];
    const double tgt = a.target_ca != nullptr ? a.target_ca[i] : a.target_value;
    const double dl[3] = { a.delta[i], a.delta[n + i], a.delta[2 * n + i] };
    const double caller_input = p.input_mode == 0 ? a.input[i] : 0.0;
    const bool synaptic = a.acc_cur != nullptr;
    const int32_t syn = synaptic ? a.acc_cur[i] : 0;
    if (en == 0) {
        a.fired[i] = fprev;  // the step leaves a disabled neuron's status as it was
        if (synaptic) {
            a.acc_cur[i] = 0;
            if (fprev != 0) {
                push_out_pairs(a, i);
            }
        }
        return;
    }
    const uint32_t step = a.step_base != nullptr ? *a.step_base + a.step : a.step;
    double input = p.input_mode == 0 ? caller_input
                                     : normal_input(a.net_seed[g], static_cast<uint32_t>(i - a.net_off[g]), step,
                                                    p.input_mean, p.input_stddev);
    if (synaptic) {
        // SynapticEquallyWeightedActivityInputCPU::update_local_input: the weights of the in-edges
        // whose source fired in the previous step, times the conductance.  The sum was pushed by those
        // sources; it is an integer, which the reference's FP64 accumulation of integer weights
        // yields exactly.  CombinedActivityInputCPU adds it to the background input.
        input = input + static_cast<double>(syn) * a.conductance;
        a.acc_cur[i] = 0;
    }
    const double scale = a.scale;
    const double ab = p.a * p.b;
    uint8_t spiked = 0;
    for (uint32_t s = 0; s < p.h; s++) {
        const double x_increase = (p.k1 * xv * xv) + (p.k2 * xv) + p.k3 - uv + input;
        const double u_increase = (ab * xv) - (p.a * uv);
        xv += x_increase * scale;
        uv += u_increase * scale;
        if (xv >= p.v_spike) {
            xv = p.c;
            uv += p.d;
            spiked = 1;
            break;
        }
    }
    a.fired[i] = spiked;
    a.x[i] = xv;
    a.u[i] = uv;
    if (synaptic && spiked != 0) {
        push_out_pairs(a, i);
    }
    const double tau_inv = a.tau_inv;
    const double local_beta = p.beta_ca * static_cast<double>(spiked == 1);
    for (uint32_t s = 0; s < p.h; s++) {
        c += scale * (c * tau_inv + local_beta);
    }
    a.ca[i] = c;
    for (int t = 0; t < 3; t++) {
        const double eta = a.min_ca != nullptr ? a.min_ca[t * n + i] : p.eta[t];
        const double clamped_target = tgt < eta ? eta : tgt;  // std::max(target, eta)
        const double dz = growth_curve(c, eta, clamped_target);
        const double nu = a.net_nu != nullptr ? a.net_nu[3 * static_cast<uint64_t>(g) + t] : p.nu[t];
        a.delta[t * n + i] = dl[t] + dz * nu;
    }
}

// Vacant retract ratios of the element commit: per neuron (3n, type-major) or one per type.
struct Retract {
    const double* per_neuron;
    double r[3];
};

// Element types [k0 / n, k1 / n) of every neuron: index k = type * n + neuron.
__global__ void k_elements_update(uint64_t n, uint64_t k0, uint64_t k1, double* __restrict__ grown,
                                  double* __restrict__ delta, uint32_t* __restrict__ c0, uint32_t* __restrict__ c1,
                                  uint32_t* __restrict__ c2, uint32_t* __restrict__ v0, uint32_t* __restrict__ v1,
                                  uint32_t* __restrict__ v2, uint32_t* __restrict__ deleted, Retract rr) {
#pragma clang fp contract(off)
    const uint64_t k = k0 + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (k >= k1) {
        return;
    }
    const int t = static_cast<int>(k / n);
    const uint64_t i = k - static_cast<uint64_t>(t) * n;
    uint32_t* connected = t == 0 ? c0 : (t == 1 ? c1 : c2);
    uint32_t* vacant = t == 0 ? v0 : (t == 1 ? v1 : v2);
    const double current_count = grown[k];
    const double current_delta = delta[k];
    const uint32_t connected_integral = connected[i];
    const double retract = rr.per_neuron != nullptr ? rr.per_neuron[k] : (t == 0 ? rr.r[0] : (t == 1 ? rr.r[1] : rr.r[2]));
    const double current_connected = static_cast<double>(connected_integral);
    const double current_vacant = current_count - current_connected;
    const double new_vacant = current_vacant + current_delta;
    if (new_vacant >= 0.0) {
        const double new_vacant_after_retract = (1 - retract) * new_vacant;
        grown[k] = new_vacant_after_retract + current_connected;
        delta[k] = 0.0;
        vacant[i] = static_cast<uint32_t>(new_vacant_after_retract);
        deleted[k] = 0;
        return;
    }
    if (current_count + current_delta <= 0.0) {
        grown[k] = 0.0;
        delta[k] = 0.0;
        vacant[i] = 0;
        connected[i] = 0;
        deleted[k] = connected_integral;
        return;
    }
    const double new_count = current_count + current_delta;
    const double new_connected_count = floor(new_count);
    const double num_vacant = new_count - new_connected_count;
    grown[k] = ((1 - retract) * num_vacant) + new_connected_count;
    delta[k] = 0.0;
    vacant[i] = 0;
    connected[i] = static_cast<uint32_t>(new_connected_count);
    deleted[k] = static_cast<uint32_t>(current_connected - new_connected_count);
}

}  // namespace

void neurons_free(rl_bh_ctx* c) {
    destroy_graph(c);
    Neurons& s = c->neurons;
    dfree(&s.x);
    dfree(&s.u);
    dfree(&s.ca);
    dfree(&s.target_ca);
    dfree(&s.input);
    dfree(&s.fired_buf[0]);
    dfree(&s.fired_buf[1]);
    s.fired = nullptr;
    s.parity = 0;
    dfree(&s.grown);
    dfree(&s.delta);
    dfree(&s.deleted);
    dfree(&s.net_nu);
    dfree(&s.min_ca);
    dfree(&s.retract);
    s.ready = false;
}

int neurons_set_element_params(rl_bh_ctx* c, const double* min_calcium, const double* retract_ratio) {
    Neurons& s = c->neurons;
    if (!s.ready) {
        return fail(c, RL_ESTATE, "rl_neurons_set_element_params before rl_neurons_init");
    }
    const size_t cnt = 3 * static_cast<size_t>(c->n);
    for (size_t k = 0; min_calcium != nullptr && k < cnt; k++) {
        if (!std::isfinite(min_calcium[k])) {
            return fail(c, RL_EINVAL, "rl_neurons_set_element_params: minimum calcium must be finite");
        }
    }
    for (size_t k = 0; retract_ratio != nullptr && k < cnt; k++) {
        if (!(retract_ratio[k] >= 0.0 && retract_ratio[k] <= 1.0)) {
            return fail(c, RL_EINVAL, "rl_neurons_set_element_params: retract ratios must lie in [0, 1]");
        }
    }
    destroy_graph(c);  // a captured period holds the old pointers
    dfree(&s.min_ca);
    dfree(&s.retract);
    if (min_calcium != nullptr && cnt > 0) {
        RL_TRY(dalloc(c, &s.min_ca, cnt));
        RL_HIP(c, hipMemcpyAsync(s.min_ca, min_calcium, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    if (retract_ratio != nullptr && cnt > 0) {
        RL_TRY(dalloc(c, &s.retract, cnt));
        RL_HIP(c, hipMemcpyAsync(s.retract, retract_ratio, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    RL_HIP(c, hipStreamSynchronize(c->stream));
    return RL_OK;
}

int neurons_set_network_growth(rl_bh_ctx* c, const double* nu) {
    Neurons& s = c->neurons;
    if (!s.ready) {
        return fail(c, RL_ESTATE, "rl_neurons_set_network_growth before rl_neurons_init");
    }
    destroy_graph(c);  // a captured period holds the old growth-rate pointer
    if (nu == nullptr) {
        dfree(&s.net_nu);
        return RL_OK;
    }
    const size_t cnt = 3 * static_cast<size_t>(c->num_networks);
    for (size_t k = 0; k < cnt; k++) {
        if (!std::isfinite(nu[k])) {
            return fail(c, RL_EINVAL, "rl_neurons_set_network_growth: growth rates must be finite");
        }
    }
    RL_TRY(dalloc(c, &s.net_nu, cnt));
    RL_HIP(c, hipMemcpyAsync(s.net_nu, nu, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
    RL_HIP(c, hipStreamSynchronize(c->stream));
    return RL_OK;
}

int neurons_init(rl_bh_ctx* c, const rl_step_params* p, const double* x, const double* u, const double* ca,
                 const double* target_ca, const double* grown, const uint32_t* connected) {
    const uint64_t n = c->n;
    Neurons& s = c->neurons;
    hipStream_t st = c->stream;
    neurons_free(c);
    s.p = *p;
    RL_TRY(dalloc(c, &s.x, n));
    RL_TRY(dalloc(c, &s.u, n));
    RL_TRY(dalloc(c, &s.ca, n));
    RL_TRY(dalloc(c, &s.target_ca, n));
    RL_TRY(dalloc(c, &s.input, n));
    RL_TRY(dalloc(c, &s.fired_buf[0], n));
    RL_TRY(dalloc(c, &s.fired_buf[1], n));
    s.parity = 0;
    s.fired = s.fired_buf[0];
    RL_TRY(dalloc(c, &s.grown, 3 * n));
    RL_TRY(dalloc(c, &s.delta, 3 * n));
    RL_TRY(dalloc(c, &s.deleted, 3 * n));
    auto up = [&](double* d, const double* h, uint64_t cnt) -> int {
        if (h != nullptr) {
            RL_HIP(c, hipMemcpyAsync(d, h, cnt * sizeof(double), hipMemcpyHostToDevice, st));
        } else {
            RL_HIP(c, hipMemsetAsync(d, 0, cnt * sizeof(double), st));
        }
        return RL_OK;
    };
    RL_TRY(up(s.x, x, n));
    RL_TRY(up(s.u, u, n));
    RL_TRY(up(s.ca, ca, n));
    RL_TRY(up(s.target_ca, target_ca, n));
    // the same bits everywhere (not merely equal values: -0 and +0 stay per-neuron)
    s.target_uniform = n > 0 && target_ca != nullptr &&
                       std::all_of(target_ca, target_ca + n,
                                   [&](double v) { return std::memcmp(&v, target_ca, sizeof(double)) == 0; });
    s.target_value = n > 0 && target_ca != nullptr ? target_ca[0] : 0.0;
    RL_TRY(up(s.grown, grown, 3 * n));
    RL_HIP(c, hipMemsetAsync(s.delta, 0, 3 * n * sizeof(double), st));
    RL_HIP(c, hipMemsetAsync(s.fired_buf[0], 0, n, st));
    RL_HIP(c, hipMemsetAsync(s.fired_buf[1], 0, n, st));
    for (int t = 0; t < 3; t++) {
        if (connected != nullptr) {
            RL_HIP(c, hipMemcpyAsync(c->conn[t], connected + t * n, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        }
    }
    RL_HIP(c, hipStreamSynchronize(st));
    s.ready = true;
    c->net.syn_valid = false;   // new fired status
    return RL_OK;
}

namespace {

StepArgs step_args(rl_bh_ctx* c) {
    Neurons& s = c->neurons;
    StepArgs a{};
    a.n = c->n;
    a.p = s.p;
    a.net_of = c->num_networks > 1 ? c->net_of : nullptr;   // one network: every neuron is in network 0
    a.net_off = c->net_off;
    a.net_seed = c->net_seed;
    a.net_nu = s.net_nu;
    a.min_ca = s.min_ca;
    a.step_base = nullptr;
    a.step = 0;
    a.x = s.x;
    a.u = s.u;
    a.ca = s.ca;
    a.input = s.input;
    a.fired_prev = s.fired_buf[s.parity];
    a.fired = s.fired_buf[s.parity ^ 1];
    a.target_ca = s.target_uniform ? nullptr : s.target_ca;
    a.target_value = s.target_value;
    a.delta = s.delta;
    a.enabled = c->enabled;
    a.scale = 1.0 / static_cast<double>(s.p.h);
    a.tau_inv = -1.0 / s.p.tau_ca;
    if (s.p.synaptic_input != 0) {
        a.acc_cur = c->net.acc[s.parity];
        a.acc_next = c->net.acc[s.parity ^ 1];
        a.out_off = c->net.out_off;
        a.out_tgt = c->net.out_tgt;
        a.out_w = c->net.out_w;
        a.conductance = s.p.synapse_conductance;
    }
    return a;
}

__global__ void k_copy_u8(uint64_t n, const uint8_t* __restrict__ from, uint8_t* __restrict__ to) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
        to[i] = from[i];
    }
}

Retract retract_of(const Neurons& s) {
    Retract r{};
    r.per_neuron = s.retract;
    for (int t = 0; t < 3; t++) {
        r.r[t] = s.p.retract_ratio[t];
    }
    return r;
}

int check_synaptic(rl_bh_ctx* c) {
    if (c->neurons.p.synaptic_input != 0 && !c->net.ready) {
        return fail(c, RL_ESTATE, "synaptic input needs the network graph (rl_net_init)");
    }
    return RL_OK;
}

}  // namespace

int neurons_step(rl_bh_ctx* c, uint32_t step, const double* input) {
    Neurons& s = c->neurons;
    if (!s.ready) {
        return fail(c, RL_ESTATE, "rl_neurons_step called before rl_neurons_init");
    }
    if (s.p.input_mode == 0 && input == nullptr) {
        return fail(c, RL_EINVAL, "input_mode 0 needs an input array");
    }
    const uint64_t n = c->n;
    if (n == 0) {
        return RL_OK;
    }
    RL_TRY(check_synaptic(c));
    if (s.p.synaptic_input != 0) {
        RL_TRY(syn_prepare(c));
    }
    if (s.p.input_mode == 0) {
        RL_HIP(c, hipMemcpyAsync(s.input, input, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    StepArgs a = step_args(c);
    a.step = step;
    k_step<<<grid_for(n), kBlock, 0, c->stream>>>(a);
    RL_HIP(c, hipGetLastError());
    s.parity ^= 1;
    s.fired = s.fired_buf[s.parity];
    return RL_OK;
}

// The last step's fired status into fired_buf[0], where a captured period expects it.
int normalize_fired(rl_bh_ctx* c) {
    Neurons& s = c->neurons;
    if (!s.ready || s.parity == 0) {
        return RL_OK;
    }
    k_copy_u8<<<kStreamBlocks, kBlock, 0, c->stream>>>(c->n, s.fired_buf[1], s.fired_buf[0]);
    RL_HIP(c, hipGetLastError());
    s.parity = 0;
    s.fired = s.fired_buf[0];
    c->net.syn_valid = false;   // the accumulators were those of parity 1
    return RL_OK;
}

// `steps` neuron steps d_steps[0] + k on the context's stream (counter-based input only).
int enqueue_neuron_steps(rl_bh_ctx* c, uint32_t steps) {
    Neurons& s = c->neurons;
    if (!s.ready || s.p.input_mode == 0) {
        return fail(c, RL_ESTATE, "captured neuron steps need rl_neurons_init with counter-based input");
    }
    if (c->n == 0) {
        return RL_OK;
    }
    RL_TRY(check_synaptic(c));
    if (s.parity != 0) {
        return fail(c, RL_ESTATE, "captured neuron steps start from fired_buf[0]");
    }
    // step k reads fired_buf[k & 1] and writes the other; an odd count copies the last status back,
    // so every replay starts and ends with it in fired_buf[0]
    StepArgs a = step_args(c);
    a.step_base = c->d_steps;
    for (uint32_t k = 0; k < steps; k++) {
        a.step = k;
        a.fired_prev = s.fired_buf[k & 1U];
        a.fired = s.fired_buf[(k + 1) & 1U];
        if (a.acc_cur != nullptr) {
            a.acc_cur = c->net.acc[k & 1U];
            a.acc_next = c->net.acc[(k + 1) & 1U];
        }
        k_step<<<grid_for(c->n), kBlock, 0, c->stream>>>(a);
        RL_HIP(c, hipGetLastError());
    }
    if ((steps & 1U) != 0) {
        k_copy_u8<<<kStreamBlocks, kBlock, 0, c->stream>>>(c->n, s.fired_buf[1], s.fired_buf[0]);
        RL_HIP(c, hipGetLastError());
    }
    return RL_OK;
}

// ElementBaseBase::update for the three element types.  With a network graph the types are committed
// one after the other, each followed by the deletion of the synapses its shrunken elements held
// (SynapseDeletionFinderCPU::delete_synapses: axons, then excitatory, then inhibitory dendrites, each
// seeing the disconnections of the previous ones); the deletion draws use the update index in
// d_steps[1].  Without a graph the deletions are only counted (neurons.deleted).
int enqueue_elements_commit(rl_bh_ctx* c) {
    Neurons& s = c->neurons;
    if (!s.ready) {
        return fail(c, RL_ESTATE, "rl_neurons_commit_elements called before rl_neurons_init");
    }
    const uint64_t n = c->n;
    if (n == 0) {
        return RL_OK;
    }
    if (!c->net.ready) {
        k_elements_update<<<grid_for(3 * n), kBlock, 0, c->stream>>>(n, 0, 3 * n, s.grown, s.delta, c->conn[0], c->conn[1],
                                                                      c->conn[2], c->vac[0], c->vac[1], c->vac[2], s.deleted,
                                                                      retract_of(s));
        RL_HIP(c, hipGetLastError());
        return RL_OK;
    }
    RL_TRY(fill_u32(c, reinterpret_cast<uint32_t*>(c->counters + kCtrDeleted), 0U, 4, nullptr, 0));
    for (int t = 0; t < 3; t++) {
        k_elements_update<<<grid_for(n), kBlock, 0, c->stream>>>(n, t * n, (t + 1) * n, s.grown, s.delta, c->conn[0],
                                                                  c->conn[1], c->conn[2], c->vac[0], c->vac[1], c->vac[2],
                                                                  s.deleted, retract_of(s));
        RL_HIP(c, hipGetLastError());
        RL_TRY(enqueue_net_delete(c, t));
    }
    return RL_OK;
}

int neurons_commit(rl_bh_ctx* c, uint32_t update, uint32_t* deleted) {
    if (c->neurons.ready && c->n > 0) {
        RL_TRY(set_steps(c, 0, update));
    }
    if (c->net.ready) {
        c->net.syn_valid = false;   // deletions change the pairs
    }
    RL_TRY(enqueue_elements_commit(c));
    if (deleted != nullptr && c->n > 0) {
        RL_HIP(c, hipMemcpyAsync(deleted, c->neurons.deleted, 3 * c->n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    unsigned long long short_of = 0;
    if (c->net.ready && c->n > 0) {
        RL_HIP(c, hipMemcpyAsync(&short_of, c->counters + kCtrDelShort, sizeof(short_of), hipMemcpyDeviceToHost,
                                 c->stream));
    }
    RL_HIP(c, hipStreamSynchronize(c->stream));
    if (short_of != 0) {
        // the reference throws here (RandomSynapseDeletionFinder::find_synapses_on_neuron); the
        // deletions that the graph allowed have been made
        return fail(c, RL_EINVAL, "rl_neurons_commit_elements: the network graph holds fewer synapses than the "
                                  "connected counts of the elements to delete");
    }
    return RL_OK;
}

int neurons_get(rl_bh_ctx* c, double* x, double* u, double* ca, uint8_t* fired, double* delta, double* grown,
                uint32_t* connected) {
    Neurons& s = c->neurons;
    if (!s.ready) {
        return fail(c, RL_ESTATE, "rl_neurons_get called before rl_neurons_init");
    }
    const uint64_t n = c->n;
    hipStream_t st = c->stream;
    if (x) RL_HIP(c, hipMemcpyAsync(x, s.x, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (u) RL_HIP(c, hipMemcpyAsync(u, s.u, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (ca) RL_HIP(c, hipMemcpyAsync(ca, s.ca, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (fired) RL_HIP(c, hipMemcpyAsync(fired, s.fired, n, hipMemcpyDeviceToHost, st));
    if (delta) RL_HIP(c, hipMemcpyAsync(delta, s.delta, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (grown) RL_HIP(c, hipMemcpyAsync(grown, s.grown, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (connected) {
        for (int t = 0; t < 3; t++) {
            RL_HIP(c, hipMemcpyAsync(connected + t * n, c->conn[t], n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        }
    }
    RL_HIP(c, hipStreamSynchronize(st));
    return RL_OK;
}

}  // namespace rl
