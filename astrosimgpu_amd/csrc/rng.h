// Counter-based random numbers shared by the device kernels and the host.
//
// Every draw is a pure function of (seed, stream, counter): splitmix64's
// finaliser applied to seed + stream*phi + counter*c1 (reference
// include/astrosimgpu/rng.hpp:26-33). A cell therefore regenerates its own
// draws from its index and the step number, with no generator state to carry
// or share, which is what lets the device integrate cells independently and
// still reproduce the host's stream.
#pragma once

#include <cstdint>

#include "amath.h"

namespace asg {

constexpr std::uint64_t kStreamMul = 0x9E3779B97F4A7C15ULL;
constexpr std::uint64_t kCounterMul = 0xBF58476D1CE4E5B9ULL;

// The (seed, stream) part of the hash input, seed + stream * phi (mod 2^64).
// Draws over streams s, s + d, s + 2d, ... advance it by d * phi, an add.
ASG_HD std::uint64_t rng_key(std::uint64_t seed, std::uint64_t stream) {
    return seed + stream * kStreamMul;
}

// splitmix64's finaliser of key + counter * c1.
ASG_HD std::uint64_t rng_bits_keyed(std::uint64_t key, std::uint64_t counter) {
    std::uint64_t z = key + counter * kCounterMul;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

ASG_HD std::uint64_t rng_bits(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter) {
    return rng_bits_keyed(rng_key(seed, stream), counter);
}

// Uniform on [0, 1) with 53 significant bits (reference rng.hpp:36-39).
ASG_HD double rng_uniform_keyed(std::uint64_t key, std::uint64_t counter) {
    return static_cast<double>(rng_bits_keyed(key, counter) >> 11) * (1.0 / 9007199254740992.0);
}

ASG_HD double rng_uniform(std::uint64_t seed, std::uint64_t stream, std::uint64_t counter) {
    return rng_uniform_keyed(rng_key(seed, stream), counter);
}

constexpr double kTwoPi = 6.283185307179586476925286766559;

// Stateless Box-Muller on counters 0 and 1, clamping u1 at 1e-300
// (reference rng.hpp:42-50). Used for the astrocyte noise current.
ASG_HD double rng_normal_clamped(std::uint64_t seed, std::uint64_t stream) {
    double u1 = rng_uniform(seed, stream, 0);
    if (u1 <= 1e-300) {
        u1 = 1e-300;
    }
    const double u2 = rng_uniform(seed, stream, 1);
    return asg_sqrt(-2.0 * asg_log(u1)) * asg_cos(kTwoPi * u2);
}

// One fresh generator's first normal() draw: redraws u1 while u1 <= 1e-300
// (reference rng.hpp:74-89); the paired sine variate is never consumed by a
// fresh per-step generator, so it is not computed. Advances `counter` past
// the draws used.
ASG_HD double rng_normal_fresh_keyed(std::uint64_t key, std::uint64_t& counter) {
    double u1;
    do {
        u1 = rng_uniform_keyed(key, counter++);
    } while (u1 <= 1e-300);
    const double u2 = rng_uniform_keyed(key, counter++);
    return asg_sqrt(-2.0 * asg_log(u1)) * asg_cos(kTwoPi * u2);
}

ASG_HD double rng_normal_fresh(std::uint64_t seed, std::uint64_t stream, std::uint64_t& counter) {
    return rng_normal_fresh_keyed(rng_key(seed, stream), counter);
}

// Knuth's multiplicative Poisson draw of a fresh generator (reference
// rng.hpp:107-124). `limit` must be exp(-lambda), computed once by the caller
// with asg_exp (the reference recomputes the same value on every call).
// Above lambda = 30 the reference switches to a rounded normal.
ASG_HD int rng_poisson_fresh_keyed(std::uint64_t key, double lambda, double limit) {
    if (lambda <= 0.0) {
        return 0;
    }
    std::uint64_t counter = 0;
    if (lambda > 30.0) {
        const double v = lambda + asg_sqrt(lambda) * rng_normal_fresh_keyed(key, counter);
        return v < 0.0 ? 0 : static_cast<int>(v + 0.5);
    }
    double product = rng_uniform_keyed(key, counter++);
    int count = 0;
    while (product > limit) {
        ++count;
        product *= rng_uniform_keyed(key, counter++);
    }
    return count;
}

// Knuth's loop of rng_poisson_fresh_keyed alone, for lambda <= 30 (the
// product of fresh uniforms until it is <= limit = e^-lambda; 0 when
// lambda <= 0, as limit >= 1 then).
ASG_HD int rng_poisson_knuth_keyed(std::uint64_t key, double limit) {
    std::uint64_t counter = 0;
    double product = rng_uniform_keyed(key, counter++);
    int count = 0;
    while (product > limit) {
        ++count;
        product *= rng_uniform_keyed(key, counter++);
    }
    return count;
}

ASG_HD int rng_poisson_fresh(std::uint64_t seed, std::uint64_t stream, double lambda,
                             double limit) {
    return rng_poisson_fresh_keyed(rng_key(seed, stream), lambda, limit);
}

// Stream salts of the reference's independent draws.
constexpr std::uint64_t kSaltNeuronPoisson = 0x2545F4914F6CDD1DULL;   // neuron.cpp:145
constexpr std::uint64_t kSaltNeuronNoise = 0x9E3779B185EBCA87ULL;     // neuron.cpp:159
constexpr std::uint64_t kSaltNoiseSharedExc = 0xD1B54A32D192ED03ULL;  // neuron.cpp:121
constexpr std::uint64_t kSaltNoiseSharedInh = 0xA24BAED4963EE407ULL;  // neuron.cpp:125
constexpr std::uint64_t kSaltAstroPoisson = 0xC2B2AE3D27D4EB4FULL;    // network.cpp:348
constexpr std::uint64_t kSaltAstroNoise = 0x9E3779B97F4A7C15ULL;      // astrocyte.cpp:244
constexpr std::uint64_t kStreamStride = 1000003ULL;                   // step/noise-index stride
constexpr std::uint64_t kNoiseStride = 1000033ULL;                    // neuron.cpp:159

}  // namespace asg
