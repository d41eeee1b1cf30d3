// Per-(pixel, sample) random streams; same keys, seeds and float mapping as
// the reference (include/core/rng.cuh:8-51, src/core/rng.cu:5-41) so that a
// path draws the same numbers in the same order on every implementation.
#pragma once

#include <cstdint>

#include "wpt_math.h"

namespace wpt {

enum RngStrategy : uint32_t { kXorShift32 = 0, kPcg32 = 1 };

WPT_HD uint64_t splitMix64(uint64_t value) {
    value += 0x9e3779b97f4a7c15ull;
    value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ull;
    value = (value ^ (value >> 27)) * 0x94d049bb133111ebull;
    return value ^ (value >> 31);
}

// Seed a stream.  XorShift32 keeps its 32-bit state in the low word.
WPT_HD uint64_t rngSeed(uint64_t key, uint32_t strategy) {
    if (strategy == kXorShift32) {
        uint32_t index = static_cast<uint32_t>(key);
        uint32_t seed = index * 747796405u + 2891336453u;
        return seed == 0 ? 1u : seed;
    }
    return splitMix64(key);
}

WPT_HD uint64_t pixelRngKey(uint32_t pixel, uint32_t sample, uint32_t strategy) {
    if (strategy == kXorShift32) return static_cast<uint64_t>(pixel ^ (sample * 0x9e3779b9u));
    return (static_cast<uint64_t>(sample) << 32) | pixel;
}

WPT_HD uint64_t photonRngKey(uint32_t photonIndex, uint32_t seed, uint32_t strategy) {
    if (strategy == kXorShift32) return static_cast<uint64_t>(seed ^ (photonIndex * 0x9e3779b9u));
    return (static_cast<uint64_t>(seed) << 32) | photonIndex;
}

WPT_HD uint32_t rngNextU32(uint64_t& state, uint32_t strategy) {
    if (strategy == kXorShift32) {
        uint32_t x = static_cast<uint32_t>(state);
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        state = x;
        return x;
    }
    uint64_t old = state;
    state = old * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
    uint32_t rotation = static_cast<uint32_t>(old >> 59u);
    return (xorshifted >> rotation) | (xorshifted << ((32u - rotation) & 31u));
}

// 24-bit uniform float in [0, 1).
WPT_HD float rngNextFloat(uint64_t& state, uint32_t strategy) {
    return static_cast<float>(rngNextU32(state, strategy) >> 8) * 5.9604644775390625e-8f;
}

// A stream bundled with its strategy, passed by reference through shading.
struct Rng {
    uint64_t state;
    uint32_t strategy;
    WPT_HD float next() { return rngNextFloat(state, strategy); }
};

}  // namespace wpt
