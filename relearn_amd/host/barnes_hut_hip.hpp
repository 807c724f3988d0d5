// C++ host adapter over librelearn_bh.so: the Barnes-Hut connectivity update behind the method names
// and argument meanings of RELeARN's algorithm interface (source/algorithm/Algorithm.h:48-197):
// init, disable_neurons, prepare_update_connectivity(signal types, vacant axons, vacant excitatory and
// inhibitory dendrites), update_connectivity(number_neurons) returning the created local synapses.
//
// Self-contained (standard library + include/relearn_bh.h only), so it can be compiled into the
// reference's build as the body of a `BarnesHutHIP : Algorithm` class.  Errors of the C ABI are thrown
// as relearn_hip::Error (a std::runtime_error carrying the library's message), the way the reference
// throws RelearnException from its checks.
#pragma once

#include <array>
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/relearn_bh.h"

namespace relearn_hip {

enum class SignalType : std::uint8_t { Excitatory = RL_SIGNAL_EXCITATORY, Inhibitory = RL_SIGNAL_INHIBITORY };

enum class Precision : std::int32_t { FP64 = RL_PREC_FP64, FP32Mixed = RL_PREC_FP32_MIXED };

// One created synapse (PlasticLocalSynapse of Connector.cpp:190-191): weight +1 for an excitatory,
// -1 for an inhibitory connection.
struct LocalSynapse {
    std::uint64_t target;
    std::uint64_t source;
    int weight;
};

// ConnectivityUpdateResult (Algorithm.h:37-42) at one rank: no distant synapses.
struct ConnectivityUpdateResult {
    std::uint64_t number_created_synapses = 0;
    std::vector<LocalSynapse> local_synapses;
};

struct BoundingBox {
    std::array<double, 3> min;
    std::array<double, 3> max;
};

class Error : public std::runtime_error {
public:
    explicit Error(const std::string& what)
        : std::runtime_error(what) { }
};

class BarnesHutHIP {
public:
    // BarnesHut(bounding_box, space_filling_curve, acceptance_criterion) (BarnesHut.h:55-60) with the
    // Gaussian kernel's sigma and mu (Gaussian.h:47-52) and the seed of the draw stream.  The default
    // precision is the CPU path's FP64, since the adapter stands in under the CPU name `barnes-hut`;
    // FP32-mixed is an explicit choice.  flags: RL_FLAG_* of relearn_bh.h (search kernel choice).
    BarnesHutHIP(const BoundingBox& box, double acceptance_criterion, double sigma = 750.0, double mu = 0.0,
                 std::uint64_t seed = 0, Precision precision = Precision::FP64, int device = 0,
                 std::int32_t flags = 0) {
        rl_bh_config cfg{};
        for (int d = 0; d < 3; d++) {
            cfg.box_min[d] = box.min[d];
            cfg.box_max[d] = box.max[d];
        }
        cfg.theta = acceptance_criterion;
        cfg.sigma = sigma;
        cfg.mu = mu;
        cfg.seed = seed;
        cfg.precision = static_cast<std::int32_t>(precision);
        cfg.device = device;
        cfg.flags = flags;
        const int rc = rl_bh_create(&cfg, &ctx_);
        if (rc != RL_OK) {
            throw Error(std::string("BarnesHutHIP: ") + rl_bh_last_error(nullptr));
        }
    }

    BarnesHutHIP(const BarnesHutHIP&) = delete;
    BarnesHutHIP& operator=(const BarnesHutHIP&) = delete;

    ~BarnesHutHIP() {
        if (ctx_ != nullptr) {
            (void)rl_bh_destroy(ctx_);
        }
    }

    // NeuronsExtraInfo positions and the axons' signal types (set before init).
    void set_neuron_positions(std::span<const std::array<double, 3>> positions) {
        positions_.assign(positions.begin(), positions.end());
    }

    void set_signal_types(std::span<const SignalType> signal_types) {
        signal_.resize(signal_types.size());
        for (std::size_t i = 0; i < signal_types.size(); i++) {
            signal_[i] = static_cast<std::uint8_t>(signal_types[i]);
        }
    }

    // Algorithm::init (Algorithm.h:111): builds the octree of the neurons' positions.
    void init(std::uint64_t number_neurons) {
        if (positions_.size() != number_neurons || signal_.size() != number_neurons) {
            throw Error("BarnesHutHIP::init: positions and signal types must be set for every neuron");
        }
        check(rl_bh_init(ctx_, number_neurons, positions_.empty() ? nullptr : positions_.front().data(),
                         signal_.data()));
        n_ = number_neurons;
        enabled_.assign(n_, 1);
    }

    // Algorithm::create_neurons (Algorithm.h:117; OctreeAlgorithm.h:73-85): the positions and signal
    // types set before hold the new neurons n .. n + creation_count - 1 as well; they are inserted into
    // the octree, enabled and without vacancies until the next prepare_update_connectivity.
    void create_neurons(std::uint64_t creation_count) {
        const std::uint64_t n1 = n_ + creation_count;
        if (creation_count == 0 || positions_.size() != n1 || signal_.size() != n1) {
            throw Error("BarnesHutHIP::create_neurons: positions and signal types must be set for the new neurons");
        }
        check(rl_bh_create_neurons(ctx_, creation_count, positions_[n_].data(), signal_.data() + n_, nullptr, nullptr,
                                   nullptr, nullptr, nullptr));
        n_ = n1;
        enabled_.resize(n_, 1);
    }

    // Algorithm::disable_neurons (Algorithm.h:125): disabled neurons neither search nor receive.
    void disable_neurons(std::span<const std::uint64_t> neuron_ids) {
        for (const auto id : neuron_ids) {
            if (id >= n_) {
                throw Error("BarnesHutHIP::disable_neurons: neuron id out of range");
            }
            enabled_[id] = 0;
        }
        check(rl_bh_set_enabled(ctx_, enabled_.data()));
    }

    // Algorithm::prepare_update_connectivity (Algorithm.h:136-140).
    void prepare_update_connectivity(std::span<const SignalType> signal_types, std::span<const std::uint32_t> vacant_axons,
                                     std::span<const std::uint32_t> vacant_excitatory_dendrites,
                                     std::span<const std::uint32_t> vacant_inhibitory_dendrites) {
        if (signal_types.size() != n_ || vacant_axons.size() != n_ || vacant_excitatory_dendrites.size() != n_
            || vacant_inhibitory_dendrites.size() != n_) {
            throw Error("BarnesHutHIP::prepare_update_connectivity: span sizes differ from the number of neurons");
        }
        for (std::size_t i = 0; i < n_; i++) {
            if (static_cast<std::uint8_t>(signal_types[i]) != signal_[i]) {
                throw Error("BarnesHutHIP::prepare_update_connectivity: signal types changed since init");
            }
        }
        check(rl_bh_prepare_update_connectivity(ctx_, vacant_axons.data(), vacant_excitatory_dendrites.data(),
                                                vacant_inhibitory_dendrites.data()));
    }

    // Algorithm::update_connectivity (Algorithm.h:149): search, conflict resolution and commit; the
    // device vacancies are updated, the created synapses returned.  The caller's synaptic elements still
    // hold the vacancies of the prepare: ForwardConnector connects one dendrite of the target (of the
    // type of the source's signal) and one axon of the source per created synapse (Connector.cpp:105,
    // 182), which the caller does with the returned synapses (weight +1: excitatory dendrite, -1:
    // inhibitory), so that the vacancies of its next prepare equal get_vacancies.
    ConnectivityUpdateResult update_connectivity(std::uint64_t number_neurons) {
        if (number_neurons != n_) {
            throw Error("BarnesHutHIP::update_connectivity: number of neurons differs from init");
        }
        std::uint64_t nreq = 0;
        check(rl_bh_find_target_neurons(ctx_, step_, &nreq));
        std::uint64_t created = 0;
        check(rl_bh_commit(ctx_, step_, nullptr, &created));
        ConnectivityUpdateResult res;
        res.number_created_synapses = created;
        if (nreq > 0) {
            std::vector<std::uint32_t> target(nreq), source(nreq), ordinal(nreq);
            std::vector<std::uint8_t> signal(nreq), ok(nreq);
            check(rl_bh_get_requests(ctx_, nreq, target.data(), source.data(), signal.data(), ordinal.data()));
            check(rl_bh_get_commit_result(ctx_, nreq, ok.data()));
            res.local_synapses.reserve(created);
            for (std::uint64_t r = 0; r < nreq; r++) {
                if (ok[r] != 0) {
                    res.local_synapses.push_back(
                        LocalSynapse{ target[r], source[r], signal[r] == RL_SIGNAL_INHIBITORY ? -1 : +1 });
                }
            }
        }
        step_++;
        return res;
    }

    // Device-resident vacancies after the last commit (what the caller's SynapticElements would hold).
    void get_vacancies(std::span<std::uint32_t> vacant_axons, std::span<std::uint32_t> vacant_excitatory_dendrites,
                       std::span<std::uint32_t> vacant_inhibitory_dendrites) {
        if (vacant_axons.size() != n_ || vacant_excitatory_dendrites.size() != n_ || vacant_inhibitory_dendrites.size() != n_) {
            throw Error("BarnesHutHIP::get_vacancies: span sizes differ from the number of neurons");
        }
        check(rl_bh_get_vacancies(ctx_, vacant_axons.data(), vacant_excitatory_dendrites.data(),
                                  vacant_inhibitory_dendrites.data()));
    }

    // Chooses the search launch tuning on the prepared update (after prepare_update_connectivity) by
    // timing it; the search results do not depend on it.  Returns the tuning kept.
    rl_bh_tuning autotune() {
        rl_bh_tuning t{};
        check(rl_bh_autotune(ctx_, step_, &t));
        return t;
    }

    // The algorithm is selected under the CPU path's name (relearn.cpp:313).
    static constexpr const char* get_algorithm_type() { return "barnes-hut"; }

    std::uint32_t plasticity_step() const { return step_; }

private:
    void check(int rc) const {
        if (rc != RL_OK) {
            throw Error(std::string("BarnesHutHIP: ") + rl_bh_last_error(ctx_));
        }
    }

    rl_bh_ctx* ctx_ = nullptr;
    std::uint64_t n_ = 0;
    std::uint32_t step_ = 0;
    std::vector<std::array<double, 3>> positions_;
    std::vector<std::uint8_t> signal_;
    std::vector<std::uint8_t> enabled_;
};

}  // namespace relearn_hip
