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

// Column pass over one matrix: workgroup handles columns [c0, c0 + C). `load(j1, j2)` gives
// the input at n = j1 + n1 j2; `store(j1, k2, v)` takes the twiddled result. C n2 / 8 lanes
// work (a workgroup may have more: a tile narrower than 2^TILE_LOG2 leaves the rest idle, but
// every lane reaches the barriers) and every working lane moves 8 elements, all 8 loads in
// flight before the first LDS write and all 8 twiddle lookups before the first store.
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
    const int lc = column_tile_log2<TILE_LOG2>(fs);
    const int C = 1 << lc;
    const int slots = column_slots(n2);
    const int lanes = (C * n2) >> 3;
    const bool working = tid < lanes;
    float2 v[8];
    if (working) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + i * lanes;
            v[i] = load(c0 + (e & (C - 1)), e >> lc);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + i * lanes;
            lds[(e & (C - 1)) * slots + lds_pad(e >> lc)] = v[i];
        }
    }
    __syncthreads();
    const int f = tid >> (fs.log2n2 - 3);  // n2 / 8 lanes per column transform
    const int t = tid & ((n2 >> 3) - 1);
    lds_fft<INV>(lds + (f < C ? f : 0) * slots, fs.log2n2, t, f < C, table);
    if (!working) {
        return;  // after the transform's last barrier
    }
    // The lane's 8 elements share the column j1 and step k2 by n2 / 8: their four-step factors
    // W^(j1 k2) are b s^i with b = W^(j1 k2_0) and s = W^(j1 n2 / 8), two table lookups
    // instead of eight (at most four roundings more).
    const unsigned j1 = static_cast<unsigned>(c0) + static_cast<unsigned>(tid & (C - 1));
    const unsigned k20 = static_cast<unsigned>(tid >> lc);
    const float2 b = four_step_twiddle32<INV>(fs, j1 * k20);
    const float2 s1 = four_step_twiddle32<INV>(fs, j1 * static_cast<unsigned>(n2 >> 3));
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
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int e = tid + i * lanes;
        const int k2 = e >> lc;
        const int col = e & (C - 1);
        v[i] = cmul(lds[col * slots + lds_pad(k2)], sp[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int e = tid + i * lanes;
        store(c0 + (e & (C - 1)), e >> lc, v[i]);
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
