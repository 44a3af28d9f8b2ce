// device_sampling.h — PCG32 and the scrambled Halton sampler on gfx950.
//
// Bit-exact with the reference HaltonSampler / PCGRng (src/sampling/HaltonSampler.h:20-113,
// src/sampling/PCGRng.h:21-41): the same 64-bit LCG, XSH-RR output, per-(pixel, dimension)
// scramble word, digit order and fused accumulation (the reference Release build contracts
// `result += float(d) * invBN` into one FMA).  Digits are computed in 32-bit integers because
// sampleIndex XOR scramble always fits in 32 bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_types.h"

namespace ahip {

constexpr uint64_t kPcgMult = 6364136223846793005ULL;
constexpr uint64_t kPcgInc = 3ULL;   // (seq=1 << 1) | 1, PCGRng.h:22-27

__device__ __forceinline__ uint32_t pcgOutput(uint64_t old) {
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

// State of PCGRng(seed) right after construction (state 0 -> step -> += seed -> step).
__device__ __forceinline__ uint64_t pcgSeededState(uint64_t seed) {
    return (kPcgInc + seed) * kPcgMult + kPcgInc;
}

__device__ __forceinline__ float radicalInverse(uint32_t b, float invB, uint32_t magic, uint32_t shifts,
                                                uint32_t n) {
    float invBN = 1.f, result = 0.f;
    if (shifts & 0x10000u) {   // n / b == mulhi(magic, n) >> sh for every 32-bit n
        const uint32_t sh = shifts & 0xffu;
        while (n > 0u) {
            uint32_t q = __umulhi(magic, n) >> sh;
            uint32_t d = n - q * b;
            n = q;
            invBN *= invB;
            result = fmaf((float)d, invBN, result);
        }
        return result;
    }
    const uint32_t sh1 = shifts & 0xffu, sh2 = shifts >> 8;
    while (n > 0u) {
        uint32_t t = __umulhi(magic, n);
        uint32_t q = (t + ((n - t) >> sh1)) >> sh2;       // == n / b exactly
        uint32_t d = n - q * b;
        n = q;
        invBN *= invB;
        result = fmaf((float)d, invBN, result);
    }
    return result;
}

// Value of dimension `dim` for sample `idx` of pixel (px, py) (HaltonSampler::get1D).
__device__ __forceinline__ float haltonSample(const HaltonTables& T, uint64_t seed, uint32_t px, uint32_t py,
                                              uint32_t idx, uint32_t dim) {
    if (dim < 128u) {
        uint64_t s = seed ^ ((uint64_t)px * 2654435761ULL) ^ ((uint64_t)py * 805459861ULL) ^
                     ((uint64_t)dim * 999999893ULL);
        uint32_t scramble = pcgOutput(pcgSeededState(s));
        return radicalInverse(T.prime[dim], T.invPrime[dim], T.magic[dim], T.shifts[dim], idx ^ scramble);
    }
    // Past 128 dimensions: the per-pixel PCG stream of startPixelSample (HaltonSampler.h:40, 45-48),
    // advanced once per fallback draw.
    uint64_t st = pcgSeededState(seed ^ (((uint64_t)px << 32) | (uint64_t)py));
    for (uint32_t k = 128u; k < dim; ++k) st = st * kPcgMult + kPcgInc;
    return (float)pcgOutput(st) * 2.3283064365386963e-10f;
}

}  // namespace ahip
