// Path random numbers: 6-round TEA (Zafar, Olano, Curtis, "GPU Random Numbers
// via the Tiny Encryption Algorithm", HPG 2010), as used by
// device/PrngState.metal:20-93 and bridge/PrngState.hpp:5-23.
#pragma once
#include "common.h"

namespace rm {

// 32-bit TEA hash of (v0, v1) (device/PrngState.metal:20-30)
RM_HD uint32_t tea32(uint32_t v0, uint32_t v1, int rounds = 6) {
    uint32_t sum = 0;
    for (int i = 0; i < rounds; ++i) {
        sum += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v1;
}

// [0, 1) float from the top 23 bits of the hash (device/PrngState.metal:49-56)
RM_HD float tea_float(uint32_t v0, uint32_t v1) {
    return u2f((tea32(v0, v1) >> 9) | 0x3f800000u) - 1.0f;
}

// Path RNG state: {seed = TEA(a, b), 16-bit draw counter} (bridge/PrngState.hpp:5-8)
struct Prng {
    uint32_t seed = 0;
    uint16_t index = 0;

    RM_HD Prng() {}
    RM_HD Prng(uint32_t a, uint32_t b) : seed(tea32(a, b)), index(0) {}
    RM_HD static Prng from_state(uint32_t s, uint16_t i) { Prng p; p.seed = s; p.index = i; return p; }

    RM_HD float sample() { return tea_float(seed, index++); }
    RM_HD vec2 sample2d() { const float a = sample(); const float b = sample(); return vec2(a, b); }
    RM_HD vec3 sample3d() { const float a = sample(); const float b = sample(); const float c = sample(); return vec3(a, b, c); }
    // uniform integer in [0, max) (device/PrngState.metal:75-77)
    RM_HD int sample_int(int max) { return (int)(tea32(seed, index++) % (uint32_t)max); }
};

}  // namespace rm
