// gen_common.hpp -- counter-based RMAT draw shared by the CPU and GPU generators.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <stdint.h>

#include "dcg_internal.hpp"

namespace dcg {

struct RmatParams {
    uint32_t scale;
    uint32_t ta, tb, tc;   // cumulative quadrant thresholds in 1/65536 units
    uint64_t seed;
};

inline RmatParams rmat_params(uint32_t scale, double a, double b, double c, uint64_t seed) {
    RmatParams p;
    p.scale = scale;
    p.ta = (uint32_t)std::llround(a * 65536.0);
    p.tb = (uint32_t)std::llround((a + b) * 65536.0);
    p.tc = (uint32_t)std::llround((a + b + c) * 65536.0);
    p.seed = seed;
    return p;
}

// Draw i: `scale` quadrant choices, most significant bit first; returns (u<<32)|v.
__host__ __device__ inline uint64_t rmat_draw(const RmatParams &p, uint64_t i) {
    uint32_t u = 0, v = 0;
    uint64_t h = 0;
    for (uint32_t l = 0; l < p.scale; ++l) {
        if ((l & 3u) == 0) h = mix64(p.seed ^ mix64(i * 16u + (l >> 2)));
        const uint32_t r = (uint32_t)(h & 0xFFFFu);
        h >>= 16;
        const uint32_t ub = r >= p.tb ? 1u : 0u;
        const uint32_t vb = (r >= p.ta && r < p.tb) || r >= p.tc ? 1u : 0u;
        u = (u << 1) | ub;
        v = (v << 1) | vb;
    }
    return ((uint64_t)u << 32) | v;
}

}  // namespace dcg
