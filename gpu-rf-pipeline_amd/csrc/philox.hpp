// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3", SC 2011): a keyed bijection of a 128-bit counter, so every sample of a
// synthetic capture is a pure function of (seed, pair, stream, index) and any lane can make any
// sample without state. Host and device share the code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rf {

struct Philox4 {
    uint32_t v[4];
};

__host__ __device__ __forceinline__ void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi,
                                                   uint32_t& lo)
{
    const uint64_t p = static_cast<uint64_t>(a) * b;
    hi = static_cast<uint32_t>(p >> 32);
    lo = static_cast<uint32_t>(p);
}

// Ten rounds with the published multipliers and Weyl key increments.
__host__ __device__ __forceinline__ Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t kM0 = 0xD2511F53u;
    constexpr uint32_t kM1 = 0xCD9E8D57u;
    constexpr uint32_t kW0 = 0x9E3779B9u;
    constexpr uint32_t kW1 = 0xBB67AE85u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(kM0, c.v[0], hi0, lo0);
        mulhilo32(kM1, c.v[2], hi1, lo1);
        c = Philox4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
        k0 += kW0;
        k1 += kW1;
    }
    return c;
}

// Uniform on [0, 1) with 24 bits (exact in float32).
__host__ __device__ __forceinline__ float uniform24(uint32_t x)
{
    return static_cast<float>(x >> 8) * 5.9604644775390625e-8f;  // 2^-24
}

// One CN(0, 1) sample (E|z|^2 = 1) from two 32-bit words by the Box-Muller transform:
// u1 in (0, 1], u2 in [0, 1), radius sqrt(-2 ln u1), angle 2 pi u2, each part scaled by
// 1/sqrt(2).
__device__ __forceinline__ float2 complex_normal(uint32_t x0, uint32_t x1)
{
    const float u1 = static_cast<float>((x0 >> 8) + 1u) * 5.9604644775390625e-8f;
    const float u2 = uniform24(x1);
    const float radius = sqrtf(-logf(u1));  // sqrt(-2 ln u1) / sqrt(2)
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    return make_float2(radius * c, radius * s);
}

}  // namespace rf
