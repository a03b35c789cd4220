// 4096-point transforms held in registers between three radix-16 Stockham passes (spans 1, 16
// and 256), 256 lanes with 16 values each. Used by the overlap-save FIR and by the row pass of
// the cross-correlations, whose rows are 4096 points long.
//
// A lane starts with the inputs t + 256 r (r < 16) in registers and ends with the outputs
// t + 256 r in registers, so a caller can read its inputs straight from HBM, apply a pointwise
// operation between two transforms without touching LDS, and write its outputs straight back;
// each transform makes two LDS round trips. Each round trip has its own LDS layout, chosen so
// that 32 consecutive lanes of its write and of its read land on distinct 8-byte bank pairs:
// the first (writes 16 t + r, reads t + 256 r) puts element j at slot1(j) = j + j/32, the
// second (writes 256 (t/16) + 16 r + t%16, reads t + 256 r) at slot2(j) = j + j/32 + 2 (j/64).
// With the earlier single layout j + j/16 (conflict-free for 16 lanes, not for 32) the overlap-
// save kernel waited 23 % longer on LDS instructions and ran 2-3 % slower. Every LDS access of
// a pass is one per-lane base plus a compile-time offset, so a pass keeps one address register
// instead of sixteen.
// Twiddles: per-device table [span][k][r] = exp(-2 pi i r k / (16 span)), r < 16, k < span,
// span 16 at offset 0 and span 256 at offset 256 (fft4096_stage_table()); a lane needs the 15
// factors r = 1..15 of row k = t % 16 of the span-16 table and of row k = t of the span-256
// table.
#pragma once

#include "fft_core.hpp"

namespace rf {

constexpr int k4096Log2 = 12;
constexpr int k4096 = 1 << k4096Log2;
constexpr int k4096Lanes = k4096 / 16;  // 256
constexpr int k4096Slots = k4096 + k4096 / 16;

__device__ __forceinline__ int pad16(int j)
{
    return j + (j >> 4);
}

// Factors 1..15 of a 16-entry table row w_r = exp(-2 pi i r k / N): entries 1, 2 (one 16-byte
// load), 4 and 8 from the table, the other eleven as products of those (w_3 = w_1 w_2, w_5 =
// w_1 w_4, w_6 = w_2 w_4, w_7 = w_3 w_4, w_8+r = w_r w_8; at most four roundings deep). Three
// loads of 32 bytes instead of eight of 120 per lane: the factor loads of every segment were
// the largest part of the overlap-save kernel's vector-memory traffic, and building them from
// four costs 44 more VALU instructions but made the kernel 16 % faster.
// Global-address loads, also when the caller's pointer has lost its address space.
__device__ __forceinline__ void load_factors(const float2* row, float2* w)
{
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const f32x4 global_f32x4;
    typedef __attribute__((address_space(1))) const double global_f64;
    const f32x4 v = *(global_f32x4*)(row + 1);  // address-space casts
    w[1] = make_float2(v.x, v.y);
    w[2] = make_float2(v.z, v.w);
    w[4] = __builtin_bit_cast(float2, *(global_f64*)(row + 4));
    w[8] = __builtin_bit_cast(float2, *(global_f64*)(row + 8));
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
#pragma unroll
    for (int r = 9; r < 16; ++r) {
        w[r] = cmul(w[r - 8], w[8]);
    }
}

// The 30 factors one lane applies in a transform. Loading them before a transform lets their
// latency overlap the caller's own loads, and one load serves a forward and an inverse
// transform (the inverse applies the conjugates).
struct Fft4096Twiddles {
    float2 w16[16];   // span-16 pass: exp(-2 pi i r (t % 16) / 256); entry 0 unused
    float2 w256[16];  // span-256 pass: exp(-2 pi i r t / 4096); entry 0 unused
};

__device__ __forceinline__ void fft4096_load_twiddles(int t, const float2* stages,
                                                      Fft4096Twiddles& tw)
{
    load_factors(stages + 16 * (t & 15), tw.w16);
    load_factors(stages + 256 + 16 * t, tw.w256);
}

// Span-1 pass from registers (inputs t + 256 r) into LDS at 16 t + r.
template <bool INV>
__device__ __forceinline__ void fft4096_first(float2* buf, int t, float2* u)
{
    dft16<INV>(u);
    float2* dst = buf + 16 * t + (t >> 1);   // slot1(16 t + r) = 16 t + t / 2 + r
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        dst[r] = u[r];
    }
    __syncthreads();
}

// Span-16 pass through LDS with the lane's factors in registers.
template <bool INV>
__device__ __forceinline__ void fft4096_middle(float2* buf, int t, const float2* w16)
{
    float2 u[16];
    const float2* src = buf + t + (t >> 5);   // slot1(t + 256 r) = t + t / 32 + 264 r
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        u[r] = src[264 * r];
    }
    __syncthreads();
    twiddled_dft16<INV>(u, w16);
    const int k = t & 15;
    float2* dst = buf + 17 * (t - k) + k;     // slot2(256 (t / 16) + 16 r + t % 16)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        dst[16 * r + (r >> 1) + 2 * (r >> 2)] = u[r];
    }
    __syncthreads();
}

// Span-256 pass from LDS; outputs t + 256 r stay in registers.
template <bool INV>
__device__ __forceinline__ void fft4096_last(const float2* buf, int t, float2* u,
                                             const float2* w256)
{
    const float2* src = buf + t + (t >> 5) + 2 * (t >> 6);   // slot2(t + 256 r) = ... + 272 r
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        u[r] = src[272 * r];
    }
    twiddled_dft16<INV>(u, w256);
}

// The whole transform, unnormalized, with preloaded factors: u[r] = x[t + 256 r] in,
// X[t + 256 r] out. `buf` holds k4096Slots values; the caller separates two uses of the same
// buffer with a barrier once every lane has read what the previous transform left there.
template <bool INV>
__device__ __forceinline__ void fft4096(float2* buf, int t, float2* u, const Fft4096Twiddles& tw)
{
    fft4096_first<INV>(buf, t, u);
    fft4096_middle<INV>(buf, t, tw.w16);
    fft4096_last<INV>(buf, t, u, tw.w256);
}

// The same, loading the factors first (they arrive while the first pass computes).
template <bool INV>
__device__ __forceinline__ void fft4096(float2* buf, int t, float2* u, const float2* stages)
{
    Fft4096Twiddles tw;
    fft4096_load_twiddles(t, stages, tw);
    fft4096<INV>(buf, t, u, tw);
}

}  // namespace rf
