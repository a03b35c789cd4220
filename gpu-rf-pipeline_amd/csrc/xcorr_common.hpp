// Device helpers shared by the single-pair and the batched cross-correlation kernels.
#pragma once

#include "rf_common.hpp"

namespace rf {

// Element n of the rotated, zero-padded a (src/xcorr.cu:278-282, 344-351).
__device__ __forceinline__ float2 rotated_a(const float2* __restrict__ a, int num_a, long long L,
                                            long long n)
{
    long long m = n + num_a - 1;
    if (m >= L) {
        m -= L;
    }
    return m < num_a ? a[m] : make_float2(0.0f, 0.0f);
}

__device__ __forceinline__ float2 padded_b(const float2* __restrict__ b, int num_b, long long n)
{
    return n < num_b ? b[n] : make_float2(0.0f, 0.0f);
}

// B conj(A) / L, or for GCC-PHAT (B conj(A)) * (1/L) / |B conj(A)| (0 where that is 0),
// with the reference's operation order (src/xcorr.cu:203-233).
__device__ __forceinline__ float2 weigh(float2 fa, float2 fb, float scale, int phat)
{
    const float re = fb.x * fa.x + fb.y * fa.y;
    const float im = fb.y * fa.x - fb.x * fa.y;
    if (!phat) {
        return make_float2(re * scale, im * scale);
    }
    const float mag = hypotf(re, im);
    const float w = mag > 0.0f ? scale / mag : 0.0f;
    return make_float2(re * w, im * w);
}

// Key ordering (magnitude, then the smaller index): magnitude bits (>= 0 floats order like
// their bit patterns) in the high word, ~index in the low word; the largest key wins.
__device__ __forceinline__ unsigned long long peak_key(float mag, long long idx)
{
    return (static_cast<unsigned long long>(__float_as_uint(mag)) << 32)
           | static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<unsigned>(idx));
}

}  // namespace rf
