// Kernels shared by every user of the library's FFT: batched single-workgroup transforms
// (P <= 4096) and the two passes of the four-step transform for longer ones.
//
// Four-step (Bailey 1990) with L = n1 * n2, n = j1 + n1 * j2, k = k2 + n2 * k1:
//   column pass: for each j1, an n2-point FFT over j2 (stride n1), times w_L^(j1 k2),
//                stored at j1 + n1 * k2;
//   row pass:    for each k2, an n1-point FFT over j1 (contiguous), giving X[k2 + n2 k1] at
//                k1 + n1 * k2 ("transposed order").
// The inverse runs the same passes in the opposite order with conjugate twiddles, taking the
// transposed order back to natural order. Cross-correlation never needs the spectrum in
// natural order, so it multiplies spectra in transposed order and skips the transposes.
#pragma once

#include "fft_core.hpp"

namespace rf {

// Split of a power-of-two length 4096 < L <= 2^24: n2 = max(8, L / 4096) rows of n1 points.
struct FourStep {
    int log2L;
    int log2n1;
    int log2n2;
    const float2* lo;  // exp(-2 pi i m / L), m < 4096
    const float2* hi;  // exp(-2 pi i 4096 m / L), m < L / 4096 (at least 1 entry)
};

inline FourStep make_four_step_shape(int log2L)
{
    FourStep fs{};
    fs.log2L = log2L;
    fs.log2n2 = log2L - kTwiddleLog2 > 3 ? log2L - kTwiddleLog2 : 3;
    fs.log2n1 = log2L - fs.log2n2;
    return fs;
}

// w_L^(+-m) for m < L.
template <bool INV>
__device__ __forceinline__ float2 four_step_twiddle(const FourStep& fs, long long m)
{
    const float2 w = cmul(gload(fs.lo, m & (kTwiddleSize - 1)), gload(fs.hi, m >> kTwiddleLog2));
    return INV ? make_float2(w.x, -w.y) : w;
}

// The same for m < L <= 2^24 in 32-bit arithmetic (every four-step transform here).
template <bool INV>
__device__ __forceinline__ float2 four_step_twiddle32(const FourStep& fs, unsigned m)
{
    const float2 w = cmul(gload(fs.lo, m & (kTwiddleSize - 1)), gload(fs.hi, m >> kTwiddleLog2));
    return INV ? make_float2(w.x, -w.y) : w;
}

// Columns per workgroup of the column pass: C * n2 = 2^TILE_LOG2 values (2^TILE_LOG2 / 8
// lanes; 4096 values and 512 lanes by default), or all n1 columns if there are fewer.
template <int TILE_LOG2 = kTwiddleLog2>
__host__ __device__ __forceinline__ int column_tile_log2(const FourStep& fs)
{
    const int c = TILE_LOG2 - fs.log2n2;
    return c < fs.log2n1 ? c : fs.log2n1;
}

template <int TILE_LOG2 = kTwiddleLog2>
__host__ __device__ __forceinline__ int column_tile(const FourStep& fs)
{
    return 1 << column_tile_log2<TILE_LOG2>(fs);
}

// LDS stride between the columns of a column tile: the padded column, n2 + n2/8 + 1 slots
// (odd for n2 >= 16, so the lanes that touch one row of consecutive columns fall on
// different banks). A tile of 2^T values takes C * column_slots(n2) = 2^T + 2^T / 8 + C.
__host__ __device__ constexpr int column_slots(int n2)
{
    return lds_slots(n2);
}

// Register-held Stockham passes for the column transforms. A transform of P = 2^log2P points
// (8 <= P <= 4096) is done by n8 = P / 8 lanes, lane t holding 8 values; the passes are those
// of lds_fft (radix 8 first, then one radix-4 or radix-2 pass; span p = 1, 8, 64, ...) with the
// same arithmetic, but a lane keeps its values in registers from one pass to the next and the
// LDS only carries the exchange between passes: the first pass works on the values as loaded
// and the last pass's results stay in registers. Value b of a pass of radix R = 2^LR is input
// position t + (b / R) n8 + (b mod R) P / R and, after the butterfly, output position
// (i - k) R + k + (b mod R) p with i = t + (b / R) n8, k = i mod p.
template <int LR>
__device__ __forceinline__ int stockham_in(int t, int n8, int P, int b)
{
    return t + (b >> LR) * n8 + (b & ((1 << LR) - 1)) * (P >> LR);
}

template <int LR>
__device__ __forceinline__ int stockham_out(int t, int n8, int p, int b)
{
    const int i = t + (b >> LR) * n8;
    const int k = i & (p - 1);
    return (i - k) * (1 << LR) + k + (b & ((1 << LR) - 1)) * p;
}

// The butterflies of one pass (span p) on the lane's 8 values, as lds_pass computes them.
template <bool INV, int LR>
__device__ __forceinline__ void stockham_pass_regs(float2 (&u)[8], int t, int n8, int p,
                                                   const float2* __restrict__ table)
{
    constexpr int R = 1 << LR;
    const int tw_step = kTwiddleSize / (p * R);
#pragma unroll
    for (int bi = 0; bi < (8 >> LR); ++bi) {
        const int i = t + bi * n8;
        twiddle_and_dft<INV>(u + bi * R, LR, i & (p - 1), table, tw_step);
    }
}

// The outputs of a pass of radix 2^LRP and span pp go to the column's LDS, the inputs of the
// next pass (radix 2^LR) come back. TAIL: a second barrier, so that the next exchange may
// write; the last exchange of a transform leaves that to the caller.
template <int LRP, int LR, bool TAIL>
__device__ __forceinline__ void stockham_exchange(float2* col, float2 (&u)[8], int t, int n8,
                                                  int P, int pp, bool active)
{
    if (active) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            col[lds_pad(stockham_out<LRP>(t, n8, pp, b))] = u[b];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            u[b] = col[lds_pad(stockham_in<LR>(t, n8, P, b))];
        }
    }
    if (TAIL) {
        __syncthreads();
    }
}

// stockham_exchange for a compile-time length P = 2^LOG2P after a radix-8 pass of span PP,
// with every LDS index written as one per-lane base plus a compile-time offset, which the LDS
// instructions carry in their offset field. lds_pad(j + c) = lds_pad(j) + c + c / 8 when c is
// a multiple of 8, and lds_pad(j) + c when j is and c < 8: the writes go to
// (t - k) 8 + k + b PP (k = t mod PP; PP = 1 or a multiple of 8), the reads come from
// t + c_b with every c_b a multiple of P / 8 (a multiple of 8 from P = 64 on).
template <int LR, bool TAIL, int LOG2P, int PP>
__device__ __forceinline__ void fixed_exchange(float2* col, float2 (&u)[8], int t, bool active)
{
    constexpr int P = 1 << LOG2P;
    constexpr int N8 = P >> 3;
    constexpr int R = 1 << LR;
    if (active) {
        const int k = t & (PP - 1);
        float2* dst = col + lds_pad((t - k) * 8 + k);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            dst[PP >= 8 ? b * PP + b * PP / 8 : b] = u[b];
        }
    }
    __syncthreads();
    if (active) {
        if constexpr (P >= 64) {
            const float2* src = col + lds_pad(t);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int c =(b >> LR) * N8 + (b & (R - 1)) * (P >> LR);
                u[b] = src[c + c / 8];
            }
        } else {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                u[b] = col[lds_pad(stockham_in<LR>(t, N8, P, b))];
            }
        }
    }
    if (TAIL) {
        __syncthreads();
    }
}

// Passes PASS .. of stockham_column_fixed (span 8^PASS onwards); returns the radix log2 of the
// last pass.
template <bool INV, int LOG2P, int PASS>
__device__ __forceinline__ int fixed_passes(float2* col, float2 (&u)[8], int t, bool active,
                                            const float2* __restrict__ table)
{
    constexpr int N8 = (1 << LOG2P) >> 3;
    constexpr int FULL8 = LOG2P / 3;
    constexpr int REM = LOG2P - 3 * FULL8;
    constexpr int p = 1 << (3 * PASS);
    if constexpr (PASS < FULL8) {
        fixed_exchange<3, (PASS + 1 < FULL8 || REM != 0), LOG2P, (p >> 3)>(col, u, t, active);
        if (active) {
            stockham_pass_regs<INV, 3>(u, t, N8, p, table);
        }
        return fixed_passes<INV, LOG2P, PASS + 1>(col, u, t, active, table);
    } else if constexpr (REM != 0) {
        fixed_exchange<REM, false, LOG2P, (p >> 3)>(col, u, t, active);
        if (active) {
            stockham_pass_regs<INV, REM>(u, t, N8, p, table);
        }
        return REM;
    } else {
        return 3;
    }
}

// stockham_column for a compile-time length: the same passes, exchanges and arithmetic (so the
// same results, bit for bit), with the index arithmetic folded into constants.
template <bool INV, int LOG2P>
__device__ __forceinline__ int stockham_column_fixed(float2* col, float2 (&u)[8], int t,
                                                     bool active, const float2* __restrict__ table)
{
    if (active) {
        stockham_pass_regs<INV, 3>(u, t, (1 << LOG2P) >> 3, 1, table);
    }
    return fixed_passes<INV, LOG2P, 1>(col, u, t, active, table);
}

// The transform of one column: on entry u[r] holds position t + n8 r, on exit u[b] holds
// output position t + n8 stockham_order(lr, b), where lr (returned) is the radix log2 of the
// last pass. Every lane of the workgroup calls it (it has barriers); `col` is the column's
// padded LDS (lds_slots(P) values) and is only touched by active lanes. After it returns, the
// caller puts a barrier before the LDS is written again. Each length 8 .. 4096 runs its own
// compile-time instance (stockham_column_fixed): with the length known only at run time, the
// LDS indices of the exchanges took a handful of instructions per access, and the column passes
// spent about as many instructions on indices as on arithmetic.
template <bool INV>
__device__ __forceinline__ int stockham_column(float2* col, float2 (&u)[8], int log2P, int t,
                                               bool active, const float2* __restrict__ table)
{
    switch (log2P) {
    case 3: return stockham_column_fixed<INV, 3>(col, u, t, active, table);
    case 4: return stockham_column_fixed<INV, 4>(col, u, t, active, table);
    case 5: return stockham_column_fixed<INV, 5>(col, u, t, active, table);
    case 6: return stockham_column_fixed<INV, 6>(col, u, t, active, table);
    case 7: return stockham_column_fixed<INV, 7>(col, u, t, active, table);
    case 8: return stockham_column_fixed<INV, 8>(col, u, t, active, table);
    case 9: return stockham_column_fixed<INV, 9>(col, u, t, active, table);
    case 10: return stockham_column_fixed<INV, 10>(col, u, t, active, table);
    case 11: return stockham_column_fixed<INV, 11>(col, u, t, active, table);
    case 12: return stockham_column_fixed<INV, 12>(col, u, t, active, table);
    default: break;
    }
    const int P = 1 << log2P;
    const int n8 = P >> 3;
    const int full8 = log2P / 3;
    const int rem = log2P - 3 * full8;
    if (active) {
        stockham_pass_regs<INV, 3>(u, t, n8, 1, table);
    }
    int p = 8;
    for (int pass = 1; pass < full8; ++pass) {
        if (pass + 1 < full8 || rem != 0) {
            stockham_exchange<3, 3, true>(col, u, t, n8, P, p >> 3, active);
        } else {
            stockham_exchange<3, 3, false>(col, u, t, n8, P, p >> 3, active);
        }
        if (active) {
            stockham_pass_regs<INV, 3>(u, t, n8, p, table);
        }
        p <<= 3;
    }
    if (rem == 2) {
        stockham_exchange<3, 2, false>(col, u, t, n8, P, p >> 3, active);
        if (active) {
            stockham_pass_regs<INV, 2>(u, t, n8, p, table);
        }
        return 2;
    }
    if (rem == 1) {
        stockham_exchange<3, 1, false>(col, u, t, n8, P, p >> 3, active);
        if (active) {
            stockham_pass_regs<INV, 1>(u, t, n8, p, table);
        }
        return 1;
    }
    return 3;
}

// Which multiple of n8 past t value b holds after a last pass of radix 2^lr (a permutation of
// 0..7): (b / R) + (b mod R) (8 / R).
__device__ __forceinline__ int stockham_order(int lr, int b)
{
    return (b >> lr) + (b & ((1 << lr) - 1)) * (8 >> lr);
}

// Column pass over one matrix: workgroup handles columns [c0, c0 + C). `load(j1, j2)` gives
// the input at n = j1 + n1 j2; `store(j1, k2, v)` takes the twiddled result. C n2 / 8 lanes
// work (a workgroup may have more: a tile narrower than 2^TILE_LOG2 leaves the rest idle, but
// every lane reaches the barriers). Lane tid takes column tid mod C and rows t + (n2 / 8) r
// (t = tid / C), so a load or store instruction covers consecutive columns of one row and every
// LDS access of the transform steps by the (odd) column stride across lanes; the transform runs
// from registers to registers (stockham_column): one LDS exchange per pass after the first,
// where staging the tile through LDS for every pass took four accesses per value more.
template <bool INV, int TILE_LOG2 = kTwiddleLog2, class Load, class Store>
__device__ __forceinline__ void column_pass_block(const FourStep& fs, long long c0, Load load,
                                                  Store store, float2* lds,
                                                  const float2* __restrict__ table)
{
    // The lane index through an empty asm: callers loop over items, and without it the
    // compiler hoists the per-lane index arithmetic of every element out of their loops into
    // held registers (123 instead of 60 for the batched forward pass).
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int n2 = 1 << fs.log2n2;
    const int n8 = n2 >> 3;
    const int lc = column_tile_log2<TILE_LOG2>(fs);
    const int C = 1 << lc;
    const int slots = column_slots(n2);
    const int lanes = (C * n2) >> 3;
    const bool working = tid < lanes;
    const int col = tid & (C - 1);
    const int t = tid >> lc;
    const long long j1l = c0 + col;
    float2 v[8];
    if (working) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            v[r] = load(j1l, t + r * n8);
        }
    }
    const int lr = stockham_column<INV>(lds + col * slots, v, fs.log2n2, t, working, table);
    if (!working) {
        return;
    }
    // The lane's 8 elements share the column j1 and step k2 by n2 / 8: their four-step factors
    // W^(j1 k2) are b s^i with b = W^(j1 k2_0) and s = W^(j1 n2 / 8), two table lookups
    // instead of eight (at most four roundings more).
    const unsigned j1 = static_cast<unsigned>(c0) + static_cast<unsigned>(col);
    const unsigned k20 = static_cast<unsigned>(t);
    const float2 b = four_step_twiddle32<INV>(fs, j1 * k20);
    const float2 s1 = four_step_twiddle32<INV>(fs, j1 * static_cast<unsigned>(n8));
    float2 sp[8];
    sp[0] = b;
    sp[1] = cmul(b, s1);
    const float2 s2 = cmul(s1, s1);
    const float2 s4 = cmul(s2, s2);
    sp[2] = cmul(b, s2);
    sp[3] = cmul(sp[1], s2);
    sp[4] = cmul(b, s4);
    sp[5] = cmul(sp[1], s4);
    sp[6] = cmul(sp[2], s4);
    sp[7] = cmul(sp[3], s4);
    if (lr == 3) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            store(j1l, t + i * n8, cmul(v[i], sp[i]));
        }
    } else if (lr == 2) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = stockham_order(2, q);
            store(j1l, t + i * n8, cmul(v[q], sp[i]));
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = stockham_order(1, q);
            store(j1l, t + i * n8, cmul(v[q], sp[i]));
        }
    }
}

// Batched single-workgroup transform: F = max(1, 512 / (P / 8)) transforms of P = 2^log2P
// (8 <= P <= 4096) points per workgroup, contiguous in memory, 512 lanes at most.
__host__ __device__ __forceinline__ int lds_batch_per_block(int log2P)
{
    const int per = 1 << (log2P - 3);
    return per >= 512 ? 1 : 512 / per;
}

template <bool INV>
__global__ __launch_bounds__(512) void fft_lds_batch_kernel(const float2* in, float2* out,
                                                            int log2P, long long batch,
                                                            float scale,
                                                            const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const int P = 1 << log2P;
    const int F = lds_batch_per_block(log2P);
    const int slots = lds_slots(P);
    const long long first = static_cast<long long>(blockIdx.x) * F;
    const int count = static_cast<int>(batch - first < F ? batch - first : F);
    const long long base = first * P;
    const int total = count * P;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int f = e >> log2P;
        lds[f * slots + lds_pad(e & (P - 1))] = in[base + e];
    }
    __syncthreads();
    const int per = P >> 3;
    const int f = threadIdx.x / per;
    const int t = threadIdx.x - f * per;
    lds_fft<INV>(lds + (f < count ? f : 0) * slots, log2P, t, f < count, table);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int fe = e >> log2P;
        out[base + e] = cscale(lds[fe * slots + lds_pad(e & (P - 1))], scale);
    }
}

}  // namespace rf
