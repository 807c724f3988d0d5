// Runs plasticity updates through the C++ adapter (barnes_hut_hip.hpp) on a network read from a raw
// input file and writes the created synapses, so that callers can check the adapter against the
// library's other bindings on the same inputs.
//
// Input file (little endian): u64 n, u32 steps, f64 theta, f64 sigma, u64 seed, i32 precision,
// f64 box_min[3], f64 box_max[3], f64 xyz[3n] (AoS), u8 signal[n], u32 vacant_axons[n],
// u32 vacant_excitatory[n], u32 vacant_inhibitory[n].
// Output file: per step u64 count, then count x (u32 target, u32 source, i32 weight); then the final
// vacancies u32[3n].  Prints one JSON line per step with the created count.
//
// The caller keeps its own synaptic elements, as the reference's Neurons do: after every update it
// connects one dendrite of each synapse's target and one axon of its source (ForwardConnector,
// Connector.cpp:105,182), and the vacancies it hands to the next prepare are its own.  It checks them
// against the device's committed vacancies (get_vacancies) and exits with status 3 when they differ.
#include <array>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "barnes_hut_hip.hpp"

namespace {

template <typename T>
void read_into(std::ifstream& f, T* dst, std::size_t count) {
    f.read(reinterpret_cast<char*>(dst), static_cast<std::streamsize>(count * sizeof(T)));
    if (!f) {
        throw std::runtime_error("update_example: truncated input file");
    }
}

template <typename T>
void write_from(std::ofstream& f, const T* src, std::size_t count) {
    f.write(reinterpret_cast<const char*>(src), static_cast<std::streamsize>(count * sizeof(T)));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: update_example <input> <output>\n";
        return 2;
    }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        if (!in) {
            std::cerr << "update_example: cannot open " << argv[1] << "\n";
            return 2;
        }
        std::uint64_t n = 0;
        std::uint32_t steps = 0;
        double theta = 0.0, sigma = 0.0;
        std::uint64_t seed = 0;
        std::int32_t precision = 0;
        relearn_hip::BoundingBox box{};
        read_into(in, &n, 1);
        read_into(in, &steps, 1);
        read_into(in, &theta, 1);
        read_into(in, &sigma, 1);
        read_into(in, &seed, 1);
        read_into(in, &precision, 1);
        read_into(in, box.min.data(), 3);
        read_into(in, box.max.data(), 3);
        std::vector<std::array<double, 3>> xyz(n);
        read_into(in, xyz.data()->data(), 3 * n);
        std::vector<std::uint8_t> sig_raw(n);
        read_into(in, sig_raw.data(), n);
        std::vector<std::uint32_t> va(n), ve(n), vi(n);
        read_into(in, va.data(), n);
        read_into(in, ve.data(), n);
        read_into(in, vi.data(), n);
        std::vector<relearn_hip::SignalType> sig(n);
        for (std::uint64_t i = 0; i < n; i++) {
            sig[i] = sig_raw[i] != 0 ? relearn_hip::SignalType::Inhibitory : relearn_hip::SignalType::Excitatory;
        }

        relearn_hip::BarnesHutHIP bh(box, theta, sigma, 0.0, seed, static_cast<relearn_hip::Precision>(precision));
        bh.set_neuron_positions(xyz);
        bh.set_signal_types(sig);
        bh.init(n);
        std::ofstream out(argv[2], std::ios::binary);
        for (std::uint32_t s = 0; s < steps; s++) {
            bh.prepare_update_connectivity(sig, va, ve, vi);
            const relearn_hip::ConnectivityUpdateResult r = bh.update_connectivity(n);
            const std::uint64_t count = r.local_synapses.size();
            write_from(out, &count, 1);
            for (const auto& syn : r.local_synapses) {
                const std::uint32_t t = static_cast<std::uint32_t>(syn.target);
                const std::uint32_t src = static_cast<std::uint32_t>(syn.source);
                const std::int32_t w = syn.weight;
                write_from(out, &t, 1);
                write_from(out, &src, 1);
                write_from(out, &w, 1);
            }
            // SynapticElements::connect_elements for each created synapse: the target's dendrite of the
            // source's signal type, the source's axon
            for (const auto& syn : r.local_synapses) {
                std::vector<std::uint32_t>& dend = syn.weight > 0 ? ve : vi;
                if (dend[syn.target] == 0 || va[syn.source] == 0) {
                    std::cerr << "update_example: a synapse without a vacant element\n";
                    return 3;
                }
                dend[syn.target]--;
                va[syn.source]--;
            }
            std::vector<std::uint32_t> da(n), de(n), di(n);
            bh.get_vacancies(da, de, di);
            if (da != va || de != ve || di != vi) {
                std::cerr << "update_example: caller-side vacancies differ from the device's after step " << s << "\n";
                return 3;
            }
            std::printf("{\"step\": %u, \"created\": %llu}\n", s, static_cast<unsigned long long>(r.number_created_synapses));
        }
        write_from(out, va.data(), n);
        write_from(out, ve.data(), n);
        write_from(out, vi.data(), n);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
}
