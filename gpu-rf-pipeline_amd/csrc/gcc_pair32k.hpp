// GCC-PHAT of a pair whose FFT length L is 2^15 (Na + Nb - 1 in (2^14, 2^15]) in one workgroup
// of 1024 lanes, with the peak search and the sub-sample offset fused in: the pair's inputs are
// read from HBM, and only its peak (and, if asked for, its output) is written back. Shared by
// rf_gcc_batch (pairs of that length) and the learned delay search (its exact fallback).
//
// The L-point spectra split into even and odd bins, each a 16384-point transform of a folded
// channel (m < L/2, x zero past its length):
//   X[2k]     = DFT16384(x[m] + x[m + L/2])
//   X[2k + 1] = DFT16384((x[m] - x[m + L/2]) exp(-2 pi i m / L)).
// The weights W = B conj(A) (1/L) / |B conj(A)| (0 where that is 0; src/xcorr.cu:219-233) of
// each half go through an inverse 16384-point transform, E from the even and O from the odd
// bins, and the unnormalized inverse L-point transform of W is
//   c[n] = E[n] + exp(2 pi i n / L) O[n],   c[n + L/2] = E[n] - exp(2 pi i n / L) O[n].
// c is the circular correlation sum_m b[m + j] conj(a[m]) under the weights: lag j for j < Nb,
// lag j - L for j > L - Na, output index lag + Na - 1. (src/xcorr.cu:278-282 rotates a by
// Na - 1 before its transform instead, which gives the same output in index order.)
//
// Registers: the lane holds the values of the operand being transformed and as many of the
// other channel's; E is parked in a per-workgroup slot of `scratch` (16384 values, written and
// read back by the same lane) between the halves. The kernel is bound by its VALU work and
// LDS exchanges, not by HBM (a timing copy with the input loads taken out is no faster), so
// the next operand's loads are not issued ahead of a transform: held in registers beside one,
// they measured slower.
// Peak: first largest |c| in output-index order (peak_key on |c|^2); sub-sample offset from the
// magnitudes of the two neighbours as in batch_finish_kernel, f = m1 / (m0 + m1) towards the
// larger neighbour m1, in float64.
#pragma once

#include "fft16384.hpp"
#include "gcc_batch_kernels.hpp"
#include "twiddle.hpp"

#include <algorithm>

namespace rf {
namespace {

constexpr int kPair32kLog2L = 15;
constexpr int kPair32kHalf = 1 << (kPair32kLog2L - 1);  // 16384

// The same pointer, opaque to the compiler, so that a load through it is issued where it is
// written (pair32k_spectra, pair32k_fold).
template <class T>
__device__ __forceinline__ T* opaque(T* p)
{
    asm volatile("" : "+s"(p));
    return p;
}

// Channel x of n samples as a buffer resource whose range check returns 0 past the end (the
// zero padding) for free: a load is one per-lane offset register (8 t) plus a scalar offset,
// with no 64-bit address or bounds test per value.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t channel_buffer(const float2* x, int n)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(x), static_cast<short>(0),
                                             n * static_cast<int>(sizeof(float2)), 0x00020000);
}

__device__ __forceinline__ float2 buffer_value(__amdgpu_buffer_rsrc_t x, int t, int m0)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                          x, static_cast<unsigned>(t) * 8u,
                                          static_cast<unsigned>(m0) * 8u, 0));
}

// Low half x[m], m = t + P j (j < V = 16384 / P), of a channel of n samples (0 past the end).
template <int P>
__device__ __forceinline__ void pair32k_load_low(const float2* __restrict__ x, int n, int t,
                                                 float2* u)
{
    const __amdgpu_buffer_rsrc_t rx = channel_buffer(x, n);
#pragma unroll
    for (int j = 0; j < kPair32kHalf / P; ++j) {
        u[j] = buffer_value(rx, t, j * P);
    }
}

// Folds the high half x[m + L/2] into u: even bins u + high, odd bins (u - high) times the
// factor exp(-2 pi i m / L) = fold_t * table[kFoldRow + j P / 512].
template <bool ODD, int P>
__device__ __forceinline__ void pair32k_fold(const float2* __restrict__ x, int n, int t,
                                             float2* u, float2 fold_t,
                                             const float2* __restrict__ table)
{
    if (n > kPair32kHalf) {
        // opaque inside the branch: range-checked buffer loads are safe to speculate, and
        // hoisted out of it they would occupy 2 V registers whether the branch runs or not
        const __amdgpu_buffer_rsrc_t rx = channel_buffer(opaque(x), n);
#pragma unroll
        for (int j = 0; j < kPair32kHalf / P; ++j) {
            const float2 h = buffer_value(rx, t, kPair32kHalf + j * P);
            u[j] = ODD ? csub(u[j], h) : cadd(u[j], h);
        }
    }
    if (ODD) {
#pragma unroll
        for (int j = 0; j < kPair32kHalf / P; ++j) {
            u[j] = cmul(u[j], cmul(fold_t, table[kFoldRow + j * (P / 512)]));
        }
    }
}

// A (into u) and B (into v) of one half of pair d's spectra: the even bins, or the odd ones.
// Every input load goes through opaque() here, where it is meant to be issued: the inputs are
// never written by the kernels, so the compiler would otherwise move the loads across the
// barriers (ahead of a transform) or keep a first load's registers for the second half. The
// caller puts a barrier before the call once every lane has read `buf`.
template <bool ODD, int P, bool SPLIT = true>
__device__ __forceinline__ void pair32k_spectra(const PairDesc& d, int t, float2* u, float2* v,
                                                float2* buf, const float2* tw, float2 fold_t,
                                                const float2* __restrict__ table)
{
    pair32k_load_low<P>(opaque(d.a), d.num_a, t, u);
    pair32k_fold<ODD, P>(opaque(d.a), d.num_a, t, u, fold_t, table);
    fft16384<false, P, SPLIT>(buf, tw, t, u);
    pair32k_load_low<P>(opaque(d.b), d.num_b, t, v);
    pair32k_fold<ODD, P>(opaque(d.b), d.num_b, t, v, fold_t, table);
    __syncthreads();
    fft16384<false, P, SPLIT>(buf, tw, t, v);
}

// One workgroup of P lanes per pair, pairs blockIdx.x, blockIdx.x + gridDim.x, ... of the list
// (or of 0 .. count - 1).
template <class R, int P>
__global__ __launch_bounds__(P) void pair32k_gcc_kernel(
    const PairDesc* __restrict__ pairs, const int* __restrict__ list,
    const int* __restrict__ list_count, int count, R* __restrict__ results,
    float2* __restrict__ scratch, const float2* __restrict__ table)
{
    constexpr int V = kPair32kHalf / P;  // values per lane and operand
    __shared__ float2 buf[k16384Slots];
    __shared__ float2 tw[kTwLds];
    __shared__ unsigned long long wave_key[P / 64];
    const int lane = threadIdx.x;
    fft16384_load_tables(tw, table, lane, P);
    const float2 fold_t = table[kFoldLane + lane];
    // E's slot as an unchecked buffer: one per-lane offset register for all its accesses
    const __amdgpu_buffer_rsrc_t keep = __builtin_amdgcn_make_buffer_rsrc(
        scratch + static_cast<long long>(blockIdx.x) * kPair32kHalf, static_cast<short>(0), -1,
        0x00020000);
    const int total = list_count != nullptr ? *list_count : count;
    const float scale = 1.0f / static_cast<float>(1 << kPair32kLog2L);
    float* power = reinterpret_cast<float*>(buf);  // 2^15 values |c|^2 after the last transform
    float2 u[V];
    float2 v[V];
    __syncthreads();  // tw
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        // The lane index through an empty asm in every iteration: otherwise the compiler hoists
        // every per-lane address of the body out of the loop into held registers.
        int t = lane;
        asm volatile("" : "+v"(t));
        const int p = list != nullptr ? list[item] : item;
        const PairDesc d = pairs[p];
        // even bins: A, B -> W -> E, parked in the workgroup's slot
        pair32k_spectra<false, P>(d, t, u, v, buf, tw, fold_t, table);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            v[j] = phat_weight(u[j], v[j], scale);
        }
        __syncthreads();
        fft16384<true, P>(buf, tw, t, v);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            using u32x2 = unsigned __attribute__((ext_vector_type(2)));
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[j]), keep,
                                                  static_cast<unsigned>(t) * 8u,
                                                  static_cast<unsigned>(j * P) * 8u, 0);
        }
        // odd bins: A, B -> W -> O
        __syncthreads();
        pair32k_spectra<true, P>(d, t, u, v, buf, tw, fold_t, table);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            v[j] = phat_weight(u[j], v[j], scale);
        }
        __syncthreads();
        fft16384<true, P>(buf, tw, t, v);
        // E back (this lane's own stores)
        const __amdgpu_buffer_rsrc_t back = __builtin_amdgcn_make_buffer_rsrc(
            opaque(scratch) + static_cast<long long>(blockIdx.x) * kPair32kHalf,
            static_cast<short>(0), -1, 0x00020000);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            u[j] = buffer_value(back, t, j * P);
        }
        __syncthreads();  // every lane has read the last transform from buf
        const int num_a = d.num_a;
        const int num_b = d.num_b;
        constexpr int L = 1 << kPair32kLog2L;
        unsigned long long best = 0ULL;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int n = t + j * P;
            const float2 f = cmul(fold_t, table[kFoldRow + j * (P / 512)]);  // exp(-2 pi i n / L)
            const float2 z = cmulconj(v[j], f);                             // O[n] exp(+2 pi i n / L)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float2 c = h == 0 ? cadd(u[j], z) : csub(u[j], z);
                const int jc = n + h * kPair32kHalf;   // circular index
                // |c|^2 orders the lags like |c| (ties aside); the square roots are taken for
                // the peak and its neighbours only
                const float pw = c.x * c.x + c.y * c.y;
                power[jc] = pw;
                const int lag = jc < num_b ? jc : (jc > L - num_a ? jc - L : L);
                if (lag != L) {
                    const unsigned long long key = peak_key(pw, lag + num_a - 1);
                    best = key > best ? key : best;
                }
            }
        }
        if (d.r != nullptr) {
            // the full output, in output-index order: c[jc] at lag + Na - 1
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int n = t + j * P;
                const float2 f = cmul(fold_t, table[kFoldRow + j * (P / 512)]);
                const float2 z = cmulconj(v[j], f);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int jc = n + h * kPair32kHalf;
                    const int lag = jc < num_b ? jc : (jc > L - num_a ? jc - L : L);
                    if (lag != L) {
                        gstore(d.r, lag + num_a - 1, h == 0 ? cadd(u[j], z) : csub(u[j], z));
                    }
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long other = __shfl_xor(best, off, 64);
            best = other > best ? other : best;
        }
        if ((t & 63) == 0) {
            wave_key[t >> 6] = best;
        }
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < P / 64; ++w) {
                best = wave_key[w] > best ? wave_key[w] : best;
            }
            const int idx = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(best & 0xFFFFFFFFULL));
            const float mag = sqrtf(__uint_as_float(static_cast<unsigned>(best >> 32)));
            const int lags = num_a + num_b - 1;
            double frac = 0.0;
            if (idx > 0 && idx + 1 < lags) {
                const int jc = (idx - (num_a - 1)) & (L - 1);   // circular index of the peak
                const double ml = sqrt(static_cast<double>(power[(jc - 1) & (L - 1)]));
                const double mh = sqrt(static_cast<double>(power[(jc + 1) & (L - 1)]));
                const double m0 = static_cast<double>(mag);
                if (m0 > 0.0) {
                    frac = mh >= ml ? mh / (m0 + mh) : -ml / (m0 + ml);
                    frac = frac > 0.5 ? 0.5 : (frac < -0.5 ? -0.5 : frac);
                }
            }
            results[p].index = idx;
            results[p].magnitude = mag;
            results[p].frac = frac;
        }
        __syncthreads();  // buf, power and wave_key are reused by the next pair
    }
}

// Workgroups of the single-workgroup GCC-PHAT: one per CU (each takes 150 KB of LDS), never
// more than the pairs; the scratch needs pair32k_scratch_values() values.
inline int pair32k_grid(int pairs)
{
    return std::max(1, std::min(pairs, cu_count()));
}

inline long long pair32k_scratch_values()
{
    return static_cast<long long>(cu_count()) * kPair32kHalf;
}

// Pairs `list` (device, *list_count entries; NULL: pairs 0 .. count - 1) of `pairs`, all with
// L = 2^15. grid_pairs bounds the grid (count, or the largest the list can hold).
constexpr int kPair32kLanes = 1024;  // measured faster than 512 lanes with 32 values each

template <class R>
void launch_pair32k(const PairDesc* pairs, const int* list, const int* list_count, int count,
                    int grid_pairs, R* results, float2* scratch, hipStream_t stream)
{
    pair32k_gcc_kernel<R, kPair32kLanes><<<static_cast<unsigned>(pair32k_grid(grid_pairs)),
                                           kPair32kLanes, 0, stream>>>(
        pairs, list, list_count, count, results, scratch, fft16384_table());
    check_launch("pair32k_gcc_kernel");
}

}  // namespace
}  // namespace rf
