// Device helpers shared by the kernels.  All translation units are compiled with
// -ffp-contract=off so that FP64 expressions round exactly like the CPU path's.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rl {

// Three-argument std::hypot of libstdc++ (<cmath>, __hypot3): scale by the largest magnitude,
// sum the three squared quotients left to right, square root, rescale.  The reference computes
// every distance on the path this way (Vec3::calculate_2_norm, external/cpp-utility/
// data-structure/Vec3.hpp:548-551).  Division and sqrt are IEEE correctly rounded on gfx950.
__device__ __forceinline__ double hypot3_exact(double x, double y, double z) {
#pragma clang fp contract(off)
    x = fabs(x);
    y = fabs(y);
    z = fabs(z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != 0.0) {
        const double xa = x / a;
        const double ya = y / a;
        const double za = z / a;
        return a * sqrt(xa * xa + ya * ya + za * za);
    }
    return a != a ? a : 0.0;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// Each round's two 32x32 -> 64-bit products are written as 64-bit multiplies, which compile to one
// v_mad_u64_u32 each instead of a v_mul_lo_u32 + v_mul_hi_u32 pair (all quarter-rate instructions).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k.x += 0x9E3779B9U;
            k.y += 0xBB67AE85U;
        }
        const uint64_t p0 = static_cast<uint64_t>(c.x) * 0xD2511F53ULL;
        const uint64_t p1 = static_cast<uint64_t>(c.z) * 0xCD9E8D57ULL;
        c = make_uint4(static_cast<uint32_t>(p1 >> 32U) ^ c.y ^ k.x, static_cast<uint32_t>(p1),
                       static_cast<uint32_t>(p0 >> 32U) ^ c.w ^ k.y, static_cast<uint32_t>(p0));
    }
    return c;
}

// Draw for the pick at `level` of axon j of neuron `neuron` in plasticity step `step` (tag 1).
__device__ __forceinline__ uint32_t pick_draw(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t j,
                                              uint32_t level) {
    const uint4 o = philox4x32_10(make_uint4(neuron, step, (j << 8U) | (level & 0xFFU), 1U),
                                  make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U)));
    return o.x;
}

// 64-bit commit priority of request (source, j) (tag 2).
__device__ __forceinline__ uint64_t commit_priority(uint64_t seed, uint32_t source, uint32_t step, uint32_t j) {
    const uint4 o = philox4x32_10(make_uint4(source, step, j, 2U),
                                  make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U)));
    return (static_cast<uint64_t>(o.x) << 32U) | o.y;
}

// 64-bit key of synapse unit `copy` between `neuron` and `partner` (ids inside their network) when
// `neuron` deletes `type` elements (0 axon, 1 excitatory, 2 inhibitory dendrite) in plasticity update
// `step`: the deleted units are the ones with the smallest keys, a uniform random subset whatever
// the order the synapses are stored in.  The low two bits of the last counter word are 0, so the
// stream never meets the tags 1-3 of the other draws.
__device__ __forceinline__ uint64_t deletion_key(uint64_t seed, uint32_t neuron, uint32_t step, uint32_t partner,
                                                 uint32_t type, uint32_t copy) {
    const uint4 o = philox4x32_10(make_uint4(neuron, step, partner, (copy << 4U) | (type << 2U)),
                                  make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32U)));
    return (static_cast<uint64_t>(o.x) << 32U) | o.y;
}

// uniform_real_distribution(0, next) of a 32-bit engine: (k / 2^32) * (next - 0) + 0
// (boost::random::detail::generate_uniform_real; ProbabilityPicker.h:108-109).
__device__ __forceinline__ double draw_to_uniform(uint32_t k, double next) {
#pragma clang fp contract(off)
    const double numerator = static_cast<double>(k);
    return numerator / 4294967296.0 * (next - 0.0) + 0.0;
}

}  // namespace rl
