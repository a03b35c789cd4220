// Device helpers shared by the single-pair and the batched cross-correlation kernels.
#pragma once

#include "fft4096.hpp"
#include "fft_passes.hpp"
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
    return m < num_a ? gload(a, m) : make_float2(0.0f, 0.0f);
}

__device__ __forceinline__ float2 padded_b(const float2* __restrict__ b, int num_b, long long n)
{
    return n < num_b ? gload(b, n) : make_float2(0.0f, 0.0f);
}

// W = B conj(A) scale / |B conj(A)| (0 where that is 0; src/xcorr.cu:219-233), without
// branches: with z = 2^e (zr, zi) scaled so that max(|zr|, |zi|) is in [0.5, 1), 1/|z| =
// 2^-e rsq(zr^2 + zi^2), the hardware reciprocal square root (about 1 ulp) of a value in
// [0.25, 2]: within a few ulp of the reference's hypotf and division, in a fraction of the
// instructions. Every GCC-PHAT path of the library weighs with it.
__device__ __forceinline__ float2 phat_weight(float2 fa, float2 fb, float scale)
{
    const float re = fb.x * fa.x + fb.y * fa.y;
    const float im = fb.y * fa.x - fb.x * fa.y;
    const float big = fmaxf(fabsf(re), fabsf(im));
    const int e = __builtin_amdgcn_frexp_expf(big);
    const float zr = ldexpf(re, -e);
    const float zi = ldexpf(im, -e);
    float w = ldexpf(scale * __builtin_amdgcn_rsqf(zr * zr + zi * zi), -e);
    w = big > 0.0f ? w : 0.0f;
    return make_float2(re * w, im * w);
}

// B conj(A) / L (src/xcorr.cu:203-214), or for GCC-PHAT phat_weight.
__device__ __forceinline__ float2 weigh(float2 fa, float2 fb, float scale, int phat)
{
    if (phat) {
        return phat_weight(fa, fb, scale);
    }
    const float re = fb.x * fa.x + fb.y * fa.y;
    const float im = fb.y * fa.x - fb.x * fa.y;
    return make_float2(re * scale, im * scale);
}

// Key ordering (magnitude, then the smaller index): magnitude bits (>= 0 floats order like
// their bit patterns) in the high word, ~index in the low word; the largest key wins.
__device__ __forceinline__ unsigned long long peak_key(float mag, long long idx)
{
    return (static_cast<unsigned long long>(__float_as_uint(mag)) << 32)
           | static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<unsigned>(idx));
}

// Row pass of the four-step cross-correlation for rows of n1 = 4096 points, 256 lanes: row
// `row` of a's and b's column-pass outputs (wa, wb point at the row) are transformed, weighed,
// transformed back and multiplied by the conjugate four-step twiddle w_L^-(j1 row), then
// written over b's row. The caller puts a barrier between two calls that share `buf`.
__device__ __forceinline__ void xcorr_row4096(const float2* __restrict__ wa,
                                              float2* __restrict__ wb, const FourStep& fs,
                                              long long row, int phat, float scale, float2* buf,
                                              const float2* __restrict__ stages)
{
    // The lane index through an empty asm: callers loop over rows, and without it the compiler
    // hoists every per-lane address of the body out of their loops into held registers.
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    float2 ua[16];
    float2 ub[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        ua[r] = gload(wa, t + r * k4096Lanes);
        ub[r] = gload(wb, t + r * k4096Lanes);
    }
    // the 30 factors of the three transforms loaded once and held (168 VGPRs in the batched
    // row pass, still three waves per SIMD): measured a little faster than loading each pass's
    // 15 factors just before the pass
    Fft4096Twiddles tw;
    fft4096_load_twiddles(t, stages, tw);
    fft4096<false>(buf, t, ua, tw);
    __syncthreads();  // every lane has read a's last pass from buf
    fft4096<false>(buf, t, ub, tw);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        ub[r] = weigh(ua[r], ub[r], scale, phat);
    }
    __syncthreads();
    fft4096<true>(buf, t, ub, tw);
    // w_L^-(j1 row), j1 = t + 256 r, as b s^r with b = w_L^(t row) and s = w_L^(256 row) (the
    // same for every lane): four table lookups instead of 32, s^2, s^4, s^8 by squaring and
    // each factor one product from an earlier one (at most five roundings deep). j1 row < L <=
    // 2^24: 32-bit index arithmetic.
    const unsigned row32 = static_cast<unsigned>(row);
    auto lookup = [&](unsigned m) {
        return cmul(gload(fs.lo, m & (kTwiddleSize - 1)), gload(fs.hi, m >> kTwiddleLog2));
    };
    float2 w[16];
    w[0] = lookup(static_cast<unsigned>(t) * row32);
    const float2 s1 = lookup(static_cast<unsigned>(k4096Lanes) * row32);
    const float2 s2 = cmul(s1, s1);
    const float2 s4 = cmul(s2, s2);
    const float2 s8 = cmul(s4, s4);
    w[1] = cmul(w[0], s1);
    w[2] = cmul(w[0], s2);
    w[3] = cmul(w[1], s2);
#pragma unroll
    for (int r = 4; r < 8; ++r) {
        w[r] = cmul(w[r - 4], s4);
    }
#pragma unroll
    for (int r = 8; r < 16; ++r) {
        w[r] = cmul(w[r - 8], s8);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int j1 = t + r * k4096Lanes;
        gstore(wb, j1, cmulconj(ub[r], w[r]));
    }
}

}  // namespace rf
