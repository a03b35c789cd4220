// 512-point transforms held by one wavefront: 64 lanes with 8 values each, three radix-8
// Stockham passes (spans 1, 8 and 64), two LDS exchanges per transform. Used by the short-
// filter overlap-save FIR, where a wave owns one 512-sample segment at a time.
//
// Lane t starts with inputs t + 64 r (r < 8) and ends with outputs t + 64 r, so a caller reads
// its inputs straight from HBM, applies a pointwise product between two transforms in
// registers and writes its outputs straight back. The passes (fft_core.hpp's Stockham map):
//   span 1:  DFT8 over the lane's own values, outputs at 8 t + r;
//   span 8:  inputs t + 64 r times exp(-2 pi i r (t % 8) / 64), outputs at 64 (t / 8) + t % 8 + 8 r;
//   span 64: inputs t + 64 r times exp(-2 pi i r t / 512), outputs at t + 64 r.
// Each exchange has its own LDS layout, chosen so that the 32 lanes of each half-wave hit 32
// distinct 8-byte bank pairs on both its write and its read:
//   exchange 1 (write 8 t + r, read t + 64 r):                 slot1(j) = j + j / 32
//   exchange 2 (write 64 (t/8) + t%8 + 8 r, read t + 64 r):     slot2(j) = j + 8 (j / 64)
// (write 1: 8 t + t / 4 + r covers 32 residues for t < 32; read 1: t + t / 32 + 66 r; write 2:
// 72 (t / 8) + t % 8 + 8 r, residues 8 (t / 8) + t % 8; read 2: t + 72 r.)
#pragma once

#include "fft_core.hpp"

namespace rf {

constexpr int k512Log2 = 9;
constexpr int k512 = 1 << k512Log2;
constexpr int k512Lanes = k512 / 8;    // 64: one wavefront
constexpr int k512Slots = 511 + 8 * 7 + 1;  // the larger layout (exchange 2): 568

// The 14 factors one lane applies in a transform, from the 4096-entry table exp(-2 pi i m /
// 4096): exp(-2 pi i r k / 64) = table[64 r k] and exp(-2 pi i r t / 512) = table[8 r t]
// (indices below 3136 and 3528). One load serves every forward and inverse transform of a
// persistent wave (the inverse applies the conjugates).
struct Fft512Twiddles {
    float2 w8[8];   // span-8 pass: exp(-2 pi i r (t % 8) / 64); entry 0 unused
    float2 w64[8];  // span-64 pass: exp(-2 pi i r t / 512); entry 0 unused
};

__device__ __forceinline__ void fft512_load_twiddles(int t, const float2* __restrict__ table,
                                                     Fft512Twiddles& tw)
{
    const int k = t & 7;
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        tw.w8[r] = table[64 * r * k];
        tw.w64[r] = table[8 * r * t];
    }
    tw.w8[0] = make_float2(1.0f, 0.0f);
    tw.w64[0] = make_float2(1.0f, 0.0f);
}

// DFT8 of (u[0], w[1] u[1], ..., w[7] u[7]) (conj(w[r]) for the inverse).
template <bool INV>
__device__ __forceinline__ void twiddled_dft8(float2 (&u)[8], const float2 (&w)[8])
{
#pragma unroll
    for (int r = 1; r < 8; ++r) {
        u[r] = cmul(u[r], INV ? make_float2(w[r].x, -w[r].y) : w[r]);
    }
    dft8<INV>(u);
}

// The whole transform, unnormalized: u[r] = x[t + 64 r] in, X[t + 64 r] out. `buf` holds
// k512Slots values and belongs to this wave; the caller separates two uses of it with a
// barrier once every lane has read the previous transform's last exchange. The barriers order
// LDS accesses only (lds_barrier), so global loads issued before the transform stay in flight.
template <bool INV>
__device__ __forceinline__ void fft512(float2* buf, int t, float2 (&u)[8],
                                       const Fft512Twiddles& tw)
{
    dft8<INV>(u);
    {
        float2* dst = buf + 8 * t + (t >> 2);   // slot1(8 t + r)
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            dst[r] = u[r];
        }
    }
    lds_barrier();
    {
        const float2* src = buf + t + (t >> 5);   // slot1(t + 64 r) = ... + 66 r
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            u[r] = src[66 * r];
        }
    }
    lds_barrier();  // every lane has read exchange 1 before exchange 2 overwrites it
    twiddled_dft8<INV>(u, tw.w8);
    {
        float2* dst = buf + 72 * (t >> 3) + (t & 7);   // slot2(64 (t / 8) + t % 8 + 8 r)
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            dst[8 * r] = u[r];
        }
    }
    lds_barrier();
    {
        const float2* src = buf + t;   // slot2(t + 64 r) = t + 72 r
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            u[r] = src[72 * r];
        }
    }
    twiddled_dft8<INV>(u, tw.w64);
}

}  // namespace rf
