// 16384-point transforms on one workgroup of P = 1024 lanes with 16 values each (or 512 lanes
// with 32), held in registers between a radix-4 pass (span 1) and three radix-16 Stockham
// passes (spans 4, 64, 1024): three LDS round trips per transform. Used by the single-workgroup GCC-PHAT of pairs whose FFT
// length is 2^15 (two half-length transforms per channel, see gcc_pair32k.hpp).
//
// A lane starts with the inputs t + P j and ends with the outputs t + P j, so a caller loads
// from HBM straight into registers, applies pointwise work between transforms without LDS, and
// stores straight from registers.
//
// LDS slot of element j: pad16(j) = j + j / 16 (17408 slots, 139 264 bytes). Every access of a
// pass is one per-lane base plus compile-time offsets (shown for P = 1024; butterfly b of a
// lane adds 17 P b / 16 to inputs and 17 P b to outputs):
//   radix-4 outputs  4 t + 4096 q + s          -> 4 t + t / 4 + 4352 q + s          (q, s < 4)
//   pass inputs      t + 1024 r                -> pad16(t) + 1088 r
//   span-4 outputs   16 (t - k) + k + 4 r      -> 17 (t - k) + k + 4 r + r / 4       (k = t % 4)
//   span-64 outputs  16 (t - k) + k + 64 r     -> 17 (t - k) + pad16(k) + 68 r       (k = t % 64)
// and the 16 lanes of each write land on 16 different 8-byte bank pairs.
//
// Twiddles come from LDS, not HBM, so that loads a caller has in flight (the next operand) are
// never waited for by a twiddle load (vector-memory counters drain in order): the span-4 and
// span-64 factors exp(-2 pi i r k / (16 span)) as direct tables (rows of 16, the span-64 rows
// padded to 17 so that 64 lanes reading 64 rows hit distinct banks; the span-64 factors are
// built from entries 1, 2, 4 and 8 of the row, apply_row_tree), the span-1024 factors
// exp(-2 pi i r k / 16384) from products of two 128-entry tables hi[m / 128] =
// exp(-2 pi i m_hi / 128) and lo[m % 128] = exp(-2 pi i m_lo / 16384) (apply_span1024).
#pragma once

#include "fft_core.hpp"
#include "fft4096.hpp"

namespace rf {

constexpr int k16384Log2 = 14;
constexpr int k16384 = 1 << k16384Log2;
constexpr int k16384Lanes = 1024;
constexpr int k16384Slots = k16384 + k16384 / 16;

// The factor tables are laid out as the kTw* / kFold* offsets of tables.hpp.

// Copies the LDS part of the table with `lanes` lanes; the caller puts a barrier before the
// first transform.
__device__ __forceinline__ void fft16384_load_tables(float2* tw, const float2* table, int t,
                                                     int lanes)
{
    for (int i = t; i < kTwLds; i += lanes) {
        tw[i] = table[i];
    }
}

// u[r] *= factor r (r = 1..15; conjugated for the inverse) in three groups of five, each
// group's factor reads kept behind a scheduling barrier: issued all at once, the 30 factor
// values (60 registers for the two-table factors) would be live together with the 32 data
// registers, and the caller's second operand would no longer fit beside the transform.
template <bool INV>
__device__ __forceinline__ void apply_row(float2* u, const float2* row)
{
#pragma unroll
    for (int g = 0; g < 3; ++g) {
#pragma unroll
        for (int r = 1 + 5 * g; r < 6 + 5 * g; ++r) {
            const float2 w = row[r];
            u[r] = INV ? cmulconj(u[r], w) : cmul(u[r], w);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// The same for the span-1024 factors w_r = exp(-2 pi i r k / 16384), k < 1024: w_1, w_2, w_4
// and w_8 from the two 128-entry tables (w_m = hi[m / 128] lo[m % 128], m < 8192), the other
// eleven as products of those (w_3 = w_1 w_2, ..., w_15 = w_7 w_8; at most four roundings deep):
// 8 LDS reads instead of 30 for the same number of multiplications.
template <bool INV>
__device__ __forceinline__ void apply_span1024(float2* u, const float2* tw, int k)
{
    auto base = [&](int m) { return cmul(tw[kTwHi + (m >> 7)], tw[kTwLo + (m & 127)]); };
    float2 w[16];
    w[1] = base(k);
    w[2] = base(2 * k);
    w[4] = base(4 * k);
    w[8] = base(8 * k);
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
#pragma unroll
    for (int r = 9; r < 16; ++r) {
        w[r] = cmul(w[r - 8], w[8]);
    }
#pragma unroll
    for (int r = 1; r < 16; ++r) {
        u[r] = INV ? cmulconj(u[r], w[r]) : cmul(u[r], w[r]);
    }
}

// Factors 1..15 of a table row from its entries 1, 2, 4 and 8 and eleven products (as in
// apply_span1024): 4 LDS reads instead of 15 where the lanes of a wave read different rows.
template <bool INV>
__device__ __forceinline__ void apply_row_tree(float2* u, const float2* row)
{
    float2 w[16];
    w[1] = row[1];
    w[2] = row[2];
    w[4] = row[4];
    w[8] = row[8];
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[1], w[4]);
    w[6] = cmul(w[2], w[4]);
    w[7] = cmul(w[3], w[4]);
#pragma unroll
    for (int r = 9; r < 16; ++r) {
        w[r] = cmul(w[r - 8], w[8]);
    }
#pragma unroll
    for (int r = 1; r < 16; ++r) {
        u[r] = INV ? cmulconj(u[r], w[r]) : cmul(u[r], w[r]);
    }
}

// Pass inputs t + P j' (butterfly t + P b, input r: j' = b + B r) into u[j'], through the
// exchange's layout: slot(t + P j') = base(t) + STRIDE j'.
template <int P, int STRIDE>
__device__ __forceinline__ void fft16384_pass_in(const float2* src, float2* u)
{
    constexpr int B = 1024 / P;
#pragma unroll
    for (int j = 0; j < 16 * B; ++j) {
        u[j] = src[STRIDE * j];
    }
}

// The whole transform on P = 1024 or 512 lanes, unnormalized: lane t holds u[j] = x[t + P j]
// (j < V = 16384 / P) in and X[t + P j] out; each radix-16 pass runs B = 1024 / P butterflies
// t + P b per lane (inputs u[b + B r]) and the radix-4 pass 4 B. `buf` holds k16384Slots values
// and `tw` the kTwLds factors; the caller separates two uses of `buf` with a barrier once every
// lane has read what the previous transform left there.
// LDS layouts of the three exchanges (write pattern, then read pattern t + P j'), as
// base(t) + compile-time offsets. SPLIT: one layout per exchange, chosen so that 32 consecutive
// lanes of each write and read land on distinct 8-byte bank pairs (j + j/32; j + j/32 +
// 2 (j/64); no padding); otherwise j + j/16 for all three (conflict-free for 16 lanes only),
// which needs fewer live address registers: measured, SPLIT makes the exact GCC-PHAT kernel
// 7 % faster and the search's forward kernel, which spills more with it, 8 % slower.
template <bool SPLIT, int P>
struct Fft16384Layout {
    // radix-4 outputs 4 t + 4 P q + s
    static __device__ __forceinline__ int w1(int t) { return SPLIT ? 4 * t + (t >> 3) : 4 * t + (t >> 2); }
    static constexpr int kW1 = SPLIT ? 33 * P / 8 : 17 * P / 4;
    // reads of the span-4, span-64 and span-1024 passes
    static __device__ __forceinline__ int r1(int t) { return SPLIT ? t + (t >> 5) : pad16(t); }
    static constexpr int kR1 = SPLIT ? 33 * P / 32 : 17 * P / 16;
    static __device__ __forceinline__ int r2(int t) { return SPLIT ? t + (t >> 5) + 2 * (t >> 6) : pad16(t); }
    static constexpr int kR2 = 17 * P / 16;
    static __device__ __forceinline__ int r3(int t) { return SPLIT ? t : pad16(t); }
    static constexpr int kR3 = SPLIT ? P : 17 * P / 16;
    // span-4 outputs 16 (t - k) + k + 4 r + 16 P b (k = t % 4): base 17 (t - k) + k in both
    static __device__ __forceinline__ constexpr int w2(int r, int b) { return 17 * P * b + 4 * r + (SPLIT ? (r >> 3) : (r >> 2)); }
    // span-64 outputs 16 (t - k) + k + 64 r + 16 P b (k = t % 64)
    static __device__ __forceinline__ int w3(int t, int k) { return SPLIT ? 16 * (t - k) + k : 17 * (t - k) + pad16(k); }
    static __device__ __forceinline__ constexpr int w3(int r, int b, int) { return SPLIT ? 16 * P * b + 64 * r : 17 * P * b + 68 * r; }
};

template <bool INV, int P, bool SPLIT = true>
__device__ __forceinline__ void fft16384(float2* buf, const float2* tw, int t, float2* u)
{
    static_assert(P == 1024 || P == 512, "16384-point transform on 1024 or 512 lanes");
    constexpr int B = 1024 / P;
    using Lay = Fft16384Layout<SPLIT, P>;
    {
        // radix 4, span 1: butterfly t + P q (q < 4 B) takes inputs t + P q + 4096 s =
        // u[q + 4 B s] and writes outputs 4 t + 4 P q + s
        float2* dst = buf + Lay::w1(t);
#pragma unroll
        for (int q = 0; q < 4 * B; ++q) {
            float2 a = u[q];
            float2 b = u[q + 4 * B];
            float2 c = u[q + 8 * B];
            float2 d = u[q + 12 * B];
            dft4<INV>(a, b, c, d);
            dst[Lay::kW1 * q + 0] = a;
            dst[Lay::kW1 * q + 1] = b;
            dst[Lay::kW1 * q + 2] = c;
            dst[Lay::kW1 * q + 3] = d;
        }
    }
    __syncthreads();
    float2 x[16];
    {
        // span 4: k = t % 4 for both butterflies of a lane (P is a multiple of 4)
        fft16384_pass_in<P, Lay::kR1>(buf + Lay::r1(t), u);
        __syncthreads();
        const int k = t & 3;
        float2* dst = buf + 17 * (t - k) + k;
#pragma unroll
        for (int b = 0; b < B; ++b) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                x[r] = u[b + B * r];
            }
            apply_row<INV>(x, tw + kTw4 + 16 * k);
            dft16<INV>(x);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                dst[Lay::w2(r, b)] = x[r];
            }
        }
    }
    __syncthreads();
    {
        // span 64: k = t % 64 for both butterflies
        fft16384_pass_in<P, Lay::kR2>(buf + Lay::r2(t), u);
        __syncthreads();
        const int k = t & 63;
        float2* dst = buf + Lay::w3(t, k);
#pragma unroll
        for (int b = 0; b < B; ++b) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                x[r] = u[b + B * r];
            }
            apply_row_tree<INV>(x, tw + kTw64 + 17 * k);
            dft16<INV>(x);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                dst[Lay::w3(r, b, 0)] = x[r];
            }
        }
    }
    __syncthreads();
    {
        // span 1024: butterfly t + P b, outputs t + P b + 1024 r stay in u[b + B r]
        fft16384_pass_in<P, Lay::kR3>(buf + Lay::r3(t), u);
#pragma unroll
        for (int b = 0; b < B; ++b) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                x[r] = u[b + B * r];
            }
            apply_span1024<INV>(x, tw, t + P * b);
            dft16<INV>(x);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                u[b + B * r] = x[r];
            }
        }
    }
}

}  // namespace rf
