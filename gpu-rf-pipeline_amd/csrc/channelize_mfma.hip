// Polyphase channelizer with the filterbank's DFT on the matrix cores (split FP16, FP32
// accumulation): rf_channelize_mfma. The definition is the reference's
// (src/channelize.cu:3-11, tools/reference.py:128-146):
//   z[t][k] = sum_{l<T} h[l] x[tM - l] exp(+2 pi i k l / M),  x[m] = 0 outside 0..N-1,
// computed as branch sums v[t][r] = sum_p h[pM + r] x[(t - p)M - r] (l = pM + r) and
//   z[t][k] = sum_r v[t][r] exp(+2 pi i k r / M),
// the second step as a real GEMM Z[t][2k + e] = sum_{2r + d} V[t][2r + d] D[2r + d][2k + e]
// (V = (Re v, Im v) interleaved, D the 128 x 128 real form of the DFT).
//
// One workgroup = 64 frames, 256 lanes (M = 64, T <= 1024, i.e. at most 16 taps per branch):
//   1 lane (r, g) keeps the 31 samples x[qM - r] its 16 frames 16g .. 16g + 15 need, and the
//     branch's 16 taps, in registers and forms the 16 branch sums in FP32 (no LDS);
//   2 the workgroup scales V by a power of two (largest |component| in [2^13, 2^14)), splits
//     it as hi + lo in FP16 into LDS (rows of 136 halves: 8 halves of padding put the 32 rows
//     a wave reads on 16 different 16-byte bank groups);
//   3 wave w computes output columns 32w .. 32w + 31 for both 32-frame row tiles with
//     v_mfma_f32_32x32x16_f16, three products per step (Vhi Dhi + Vhi Dlo + Vlo Dhi; Vlo Dlo,
//     2^-22 of a product, is dropped); its D fragments (hi and lo, computed in float64 at
//     build time, tables.hpp) stay in registers;
//   4 the FP32 result is scaled back exactly and written as 32 consecutive floats per row.

#include "fir_internal.hpp"
#include "rf_common.hpp"
#include "twiddle.hpp"

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace rf {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kM = kDftChannels;       // channels
constexpr int kP = 16;                 // taps per branch (T <= kP * kM)
constexpr int kFramesPerBlock = 64;
constexpr int kThreads = 256;
constexpr int kRowHalves = 2 * kM + 8; // 136: one V row in LDS
constexpr int kSteps = kDftSteps;      // 8 k-steps of 16
constexpr int kTableHalves = kDftHalves;  // per hi or lo: [w][s][lane][q] (tables.hpp)

// Power-of-two exponent e with max * 2^e in [2^13, 2^14) (0 for max == 0).
__device__ __forceinline__ int scale_exponent(float max_abs)
{
    if (!(max_abs > 0.0f)) {
        return 0;
    }
    int e = 0;
    (void)frexpf(max_abs, &e);
    const int s = 14 - e;
    return s > 100 ? 100 : (s < -100 ? -100 : s);
}

__global__ __launch_bounds__(kThreads) void channelize_mfma_kernel(
    const float2* __restrict__ x, const float* __restrict__ taps, float* __restrict__ out,
    int n, int num_taps, long long frames, const _Float16* __restrict__ table)
{
    __shared__ __align__(16) _Float16 vhi[kFramesPerBlock * kRowHalves];
    __shared__ __align__(16) _Float16 vlo[kFramesPerBlock * kRowHalves];
    __shared__ float wave_max[kThreads / 64];
    const int t = threadIdx.x;
    const int l = t & 63;
    const int w = t >> 6;
    // XCD-major order: workgroup b runs on XCD b % 8, which gets consecutive frame blocks, so
    // the 15-frame history each block re-reads was just loaded into that XCD's L2.
    const long long blocks = (frames + kFramesPerBlock - 1) / kFramesPerBlock;
    const long long per_xcd = (blocks + 7) / 8;
    const long long block = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (block >= blocks) {
        return;  // the whole workgroup: padding of the grid to a multiple of 8
    }
    const long long t0 = block * kFramesPerBlock;

    // 1. branch sums: lane (r, g) = (t % 64, t / 64), frames f = 16 g + j.
    const int r = l;
    const int g = w;
    float hr[kP];
#pragma unroll
    for (int p = 0; p < kP; ++p) {
        const int k = p * kM + r;
        hr[p] = k < num_taps ? taps[k] : 0.0f;
    }
    float2 xq[kP + 15];  // x[(q0 + u) M - r], q0 = t0 + 16 g - 15
    const long long q0 = t0 + 16 * g - (kP - 1);
    if (q0 * kM - (kM - 1) >= 0 && (q0 + kP + 14) * kM < n) {
        // Every sample of the window inside x: no per-load test.
        const float2* src = x + q0 * kM - r;
#pragma unroll
        for (int u = 0; u < kP + 15; ++u) {
            xq[u] = gload(src, u * kM);
        }
    } else {
#pragma unroll
        for (int u = 0; u < kP + 15; ++u) {
            const long long m = (q0 + u) * kM - r;
            xq[u] = (m >= 0 && m < n) ? x[m] : make_float2(0.0f, 0.0f);
        }
    }
    float2 v[16];
    float vmax = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int p = 0; p < kP; ++p) {
            const float2 s = xq[j + (kP - 1) - p];
            acc.x = fmaf(hr[p], s.x, acc.x);
            acc.y = fmaf(hr[p], s.y, acc.y);
        }
        v[j] = acc;
        vmax = fmaxf(vmax, fmaxf(fabsf(acc.x), fabsf(acc.y)));
    }

    // D fragments of this wave, issued after the branch sums (whose sample window holds most
    // of the registers) and used by the GEMM after two barriers.
    f16x8 dh[kSteps];
    f16x8 dl[kSteps];
    const f16x8* tab = reinterpret_cast<const f16x8*>(table);
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        dh[s] = tab[(w * kSteps + s) * 64 + l];
        dl[s] = tab[kTableHalves / 8 + (w * kSteps + s) * 64 + l];
    }

    // 2. block scale, FP16 split into LDS rows 16 g + j, columns 2 r, 2 r + 1.
    for (int off = 32; off > 0; off >>= 1) {
        vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    }
    if (l == 0) {
        wave_max[w] = vmax;
    }
    __syncthreads();
    const float bmax = fmaxf(fmaxf(wave_max[0], wave_max[1]), fmaxf(wave_max[2], wave_max[3]));
    const int ex = scale_exponent(bmax);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float re = ldexpf(v[j].x, ex);
        const float im = ldexpf(v[j].y, ex);
        const _Float16 rh = static_cast<_Float16>(re);
        const _Float16 ih = static_cast<_Float16>(im);
        const int o = (16 * g + j) * kRowHalves + 2 * r;
        vhi[o] = rh;
        vhi[o + 1] = ih;
        vlo[o] = static_cast<_Float16>(re - static_cast<float>(rh));
        vlo[o + 1] = static_cast<_Float16>(im - static_cast<float>(ih));
    }
    __syncthreads();

    // 3. GEMM: row tiles 0 and 1 (frames 0..31, 32..63), columns 32 w .. 32 w + 31.
    const int i = l & 31;
    const int kq = 8 * (l >> 5);
    f32x16 acc0 = {};
    f32x16 acc1 = {};
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        const int c = 16 * s + kq;
        const f16x8 a0h = *reinterpret_cast<const f16x8*>(vhi + i * kRowHalves + c);
        const f16x8 a0l = *reinterpret_cast<const f16x8*>(vlo + i * kRowHalves + c);
        const f16x8 a1h = *reinterpret_cast<const f16x8*>(vhi + (32 + i) * kRowHalves + c);
        const f16x8 a1l = *reinterpret_cast<const f16x8*>(vlo + (32 + i) * kRowHalves + c);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0h, dh[s], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1h, dh[s], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0h, dl[s], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1h, dl[s], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0l, dh[s], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1l, dh[s], acc1, 0, 0, 0);
    }

    // 4. scale back, write z[t][k].{re, im} = out[2 (t M + k) + e] = out[t 128 + col].
    const float scale = ldexpf(1.0f, -ex);
    const int col = 32 * w + i;
    if (t0 + kFramesPerBlock <= frames) {
        float* dst = out + t0 * (2 * kM) + col;
        // non-temporal, as the fused kernel's stores: the outputs are not re-read here
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int row = (q & 3) + 8 * (q >> 2) + 4 * (l >> 5);
            __builtin_nontemporal_store(acc0[q] * scale, dst + row * (2 * kM));
            __builtin_nontemporal_store(acc1[q] * scale, dst + (32 + row) * (2 * kM));
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int row = (q & 3) + 8 * (q >> 2) + 4 * (l >> 5);
        const long long f0 = t0 + row;
        const long long f1 = t0 + 32 + row;
        if (f0 < frames) {
            out[f0 * (2 * kM) + col] = acc0[q] * scale;
        }
        if (f1 < frames) {
            out[f1 * (2 * kM) + col] = acc1[q] * scale;
        }
    }
}

}  // namespace
}  // namespace rf

extern "C" rf_status rf_channelize_mfma(const rf_cf32* x, const float* taps, rf_cf32* out,
                                        int num_samples, int channels, int num_taps,
                                        rf_stream stream)
{
    return rf::guarded([&] {
        if (channels < 1) rf::fail(RF_INVALID_ARG, "channelize mfma: need at least one channel");
        if (num_taps < 1) rf::fail(RF_INVALID_ARG, "channelize mfma: need at least one tap");
        if (num_samples < 0) rf::fail(RF_INVALID_ARG, "channelize mfma: negative sample count");
        if (channels != rf::kM || num_taps > rf::kP * rf::kM) {
            rf::fail(RF_LIMIT, "channelize mfma: supports " + std::to_string(rf::kM)
                                   + " channels and at most " + std::to_string(rf::kP * rf::kM)
                                   + " taps (got " + std::to_string(channels) + ", "
                                   + std::to_string(num_taps) + ")");
        }
        const long long frames = rf::ceil_div(num_samples, channels);
        if (frames == 0) {
            return;
        }
        if (x == nullptr || taps == nullptr || out == nullptr) {
            rf::fail(RF_INVALID_ARG, "channelize mfma: NULL buffer");
        }
        const _Float16* table = rf::dft64_table();
        const unsigned grid =
            static_cast<unsigned>(8 * rf::ceil_div(rf::ceil_div(frames, rf::kFramesPerBlock), 8));
        rf::channelize_mfma_kernel<<<grid, rf::kThreads, 0, rf::as_stream(stream)>>>(
            reinterpret_cast<const float2*>(x), taps, reinterpret_cast<float*>(out),
            num_samples, num_taps, frames, table);
        rf::check_launch("channelize_mfma_kernel");
    });
}
