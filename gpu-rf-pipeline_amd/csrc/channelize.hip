// Polyphase channelizer for gfx950 (rf_channelize).
//
// Definition (src/channelize.cu:3-11, tools/reference.py:24-30,128-146):
//   u[t][m] = sum_p h[pM + m] x[tM - pM - m]               (branch sums, x[<0] = 0)
//   z[t][k] = sum_m u[t][m] exp(+2 pi i k m / M)          (unnormalised inverse DFT)
//
// Fused path (M a power of two, 8 <= M <= 4096): a workgroup of 512 lanes owns 4096 / M
// consecutive frames. It stages the samples those frames reach back to in LDS, computes the
// branch sums (for M <= 512 each lane owns one branch and eight consecutive frames and slides a
// register window along the branch, reading each staged sample once), writes u[t][.] into the
// same LDS, runs the M-point inverse FFTs there and stores z with coalesced writes. x is read
// from HBM once (plus the (P-1)M-sample history per workgroup) and z written once.
// Other M: a branch-sum kernel and the plan's batched inverse FFT, as the reference does.

#include "fft_plan.hpp"
#include "twiddle.hpp"

#include <string>

namespace rf {
namespace {

__device__ __forceinline__ float2 load_sample(const float2* __restrict__ x, long long m, int n)
{
    return (m >= 0 && m < n) ? x[m] : make_float2(0.0f, 0.0f);
}

// One lane per (frame, branch), as src/channelize.cu:25-49.
__global__ void channelize_branch_kernel(const float2* __restrict__ x,
                                         const float* __restrict__ taps, float2* __restrict__ u,
                                         int channels, int num_taps, long long frames)
{
    const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= frames * channels) {
        return;
    }
    const long long t = idx / channels;
    const int m = static_cast<int>(idx - t * channels);
    const long long newest = t * channels;
    float re = 0.0f;
    float im = 0.0f;
    for (int l = m; l < num_taps; l += channels) {
        const long long s = newest - l;
        if (s < 0) {
            break;
        }
        const float h = taps[l];
        const float2 v = x[s];
        re = fmaf(h, v.x, re);
        im = fmaf(h, v.y, im);
    }
    u[idx] = make_float2(re, im);
}

constexpr int kChanThreads = 512;
constexpr int kChanOutputs = 4096;  // frames * M per workgroup
constexpr int kChanMaxWindowTaps = 16;

// LDS stride between the frames of the inverse FFTs. For M = 64 each radix-8 pass reads and
// writes slots f S + 9 r + t (t < 8 the lane within frame f): with S = 72 (8 mod 32) the four
// frames of a 32-lane group fill the 32 bank pairs once, where the padded transform's own
// stride (73) made the first pass's writes 9 (f + t) mod 32 collide. Other M keep lds_slots.
__host__ __device__ constexpr int chan_frame_slots(int log2M)
{
    return log2M == 6 ? 72 : lds_slots(1 << log2M);
}
static_assert(lds_pad(63) < 72, "a 64-point frame fits 72 slots");

// Branch sums of one lane for M <= 512: branch m, frames 8g .. 8g+7 of the workgroup.
// PB (1..16) taps per branch at compile time; lds holds x[t0 M - pb M + 1 + i]. The window of
// 8 + PB - 1 samples along the branch is streamed from the newest sample to the oldest, each
// sample going into every output that uses it, so only the eight accumulators and the taps
// stay live, and each output still sums its taps in ascending order (p = 0, 1, ...), as
// src/channelize.cu:40-48 does.
// The lane's PB taps h[pM + m] (0 past the filter), loaded before the workgroup stages its
// samples so that their latency hides behind the staging loads instead of following the
// staging barrier.
template <int PB>
__device__ __forceinline__ void branch_taps(const float* __restrict__ taps, int num_taps, int M,
                                            int m, float* h)
{
#pragma unroll
    for (int p = 0; p < PB; ++p) {
        const int l = p * M + m;
        h[p] = l < num_taps ? taps[l] : 0.0f;
    }
}

template <int PB>
__device__ __forceinline__ void branch_window(const float2* lds, const float* h, int M, int m,
                                              int g, float2* acc)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        acc[j] = make_float2(0.0f, 0.0f);
    }
    constexpr int W = 8 + PB - 1;
#pragma unroll
    for (int q = W - 1; q >= 0; --q) {
        const float2 v = lds[(8 * g + 1 + q) * M - 1 - m];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = j - q + PB - 1;  // output j uses window sample q with tap p
            if (p >= 0 && p < PB) {
                acc[j].x = fmaf(h[p], v.x, acc[j].x);
                acc[j].y = fmaf(h[p], v.y, acc[j].y);
            }
        }
    }
}

// Generic branch sum from the staged samples: output (f, m) of the workgroup.
__device__ __forceinline__ float2 branch_generic(const float2* lds, const float* __restrict__ taps,
                                                 int num_taps, int M, int pb, int f, int m)
{
    float re = 0.0f;
    float im = 0.0f;
    for (int p = 0; p < pb; ++p) {
        const int l = p * M + m;
        const float h = l < num_taps ? taps[l] : 0.0f;
        const float2 v = lds[(f - p + pb) * M - 1 - m];
        re = fmaf(h, v.x, re);
        im = fmaf(h, v.y, im);
    }
    return make_float2(re, im);
}

// LOG2M != 0 fixes the channel count at compile time (two radix-8 passes for M = 64, the
// common filterbank size), which keeps the frame FFTs in about half the registers of the
// general-length transform.
template <int PB, int LOG2M>
__global__ __launch_bounds__(kChanThreads) void channelize_fused_kernel(
    const float2* __restrict__ x, const float* __restrict__ taps, float2* __restrict__ out,
    int n, int log2M_rt, int num_taps, long long frames, const float2* __restrict__ table)
{
    // The staged samples sit one slot in (st = lds + 1), so that the interior groups' sample
    // pairs, which start one sample before the group, land on 16-byte LDS units: one 16-byte
    // LDS write per pair instead of two 8-byte writes to alternate slots (two-way bank
    // conflicts). Slot 0 and the slot after the last sample take the dropped ends.
    extern __shared__ __attribute__((aligned(16))) float2 lds[];
    float2* const st = lds + 1;
    const int log2M = LOG2M ? LOG2M : log2M_rt;
    const int M = 1 << log2M;
    const int F = kChanOutputs >> log2M;
    const int pb = PB ? PB : (num_taps + M - 1) / M;
    // XCD-major order: workgroup b runs on XCD b % 8, which gets consecutive frame groups, so
    // the (pb-1) M history each group re-reads was just loaded into that XCD's L2.
    const long long groups = (frames + F - 1) / F;
    const long long per_xcd = (groups + 7) / 8;
    const long long group = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (group >= groups) {
        return;  // the whole workgroup: padding of the grid to a multiple of 8
    }
    const long long t0 = group * F;
    constexpr bool windowed = PB != 0;
    float h[PB ? PB : 1];
    if (windowed) {
        branch_taps<PB ? PB : 1>(taps, num_taps, M, threadIdx.x & (M - 1), h);
    }
    const long long base = t0 * M - static_cast<long long>(pb) * M + 1;
    const int len = (F + pb - 1) * M;
    if constexpr (LOG2M != 0 && PB != 0) {
        // Fixed shape: every lane's loads issued before the first LDS store, so the staging
        // costs one HBM round trip instead of one per loop iteration.
        constexpr int kLen = ((kChanOutputs >> LOG2M) + PB - 1) << LOG2M;
        constexpr int kIters = (kLen + kChanThreads - 1) / kChanThreads;
        // Interior groups whose samples base - 1 .. base + kLen start 16-byte aligned load
        // them as pairs (one 16-byte load per two samples, half the load instructions; the
        // first and last sample of the pairs fall outside the group's range and land in the
        // two spare slots).
        constexpr int kPairs = kLen / 2 + 1;
        constexpr int kPairIters = (kPairs + kChanThreads - 1) / kChanThreads;
        const unsigned long long pair0 =
            reinterpret_cast<unsigned long long>(x) + static_cast<unsigned long long>(base - 1) * 8ULL;
        if (base >= 1 && base + kLen + 1 <= n && (pair0 & 15ULL) == 0) {
            typedef float f32x4 __attribute__((ext_vector_type(4)));
            typedef __attribute__((address_space(1))) const f32x4 global_f32x4;
            const global_f32x4* src = reinterpret_cast<const global_f32x4*>(pair0);
            f32x4 w[kPairIters];
#pragma unroll
            for (int k = 0; k < kPairIters; ++k) {
                const int q = threadIdx.x + k * kChanThreads;
                w[k] = q < kPairs ? src[q] : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            }
            // pair q holds samples 2 q - 1 and 2 q of the group: st[2 q - 1 .. 2 q] = lds[2 q ..]
            f32x4* const lds4 = reinterpret_cast<f32x4*>(lds);
#pragma unroll
            for (int k = 0; k < kPairIters; ++k) {
                const int q = threadIdx.x + k * kChanThreads;
                if (q < kPairs) {
                    lds4[q] = w[k];
                }
            }
        } else {
            float2 v[kIters];
            if (base >= 0 && base + kLen <= n) {
                const float2* src = x + base;
#pragma unroll
                for (int k = 0; k < kIters; ++k) {
                    const int i = threadIdx.x + k * kChanThreads;
                    v[k] = i < kLen ? gload(src, i) : make_float2(0.0f, 0.0f);
                }
            } else {
#pragma unroll
                for (int k = 0; k < kIters; ++k) {
                    const int i = threadIdx.x + k * kChanThreads;
                    v[k] = i < kLen ? load_sample(x, base + i, n) : make_float2(0.0f, 0.0f);
                }
            }
#pragma unroll
            for (int k = 0; k < kIters; ++k) {
                const int i = threadIdx.x + k * kChanThreads;
                if (i < kLen) {
                    st[i] = v[k];
                }
            }
        }
    } else {
        for (int i = threadIdx.x; i < len; i += kChanThreads) {
            st[i] = load_sample(x, base + i, n);
        }
    }
    __syncthreads();

    // PB != 0 is only launched for M <= 512 (try_fused), where each lane owns one branch.
    float2 acc[8];
    if (windowed) {
        const int m = threadIdx.x & (M - 1);
        const int g = threadIdx.x >> log2M;
        branch_window<PB ? PB : 1>(st, h, M, m, g, acc);
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = threadIdx.x + j * kChanThreads;
            acc[j] = branch_generic(st, taps, num_taps, M, pb, e >> log2M, e & (M - 1));
        }
    }
    __syncthreads();  // the staged samples are dead; the frames reuse the LDS

    const int slots = chan_frame_slots(log2M);
    if (windowed) {
        const int m = threadIdx.x & (M - 1);
        const int g = threadIdx.x >> log2M;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            lds[(8 * g + j) * slots + lds_pad(m)] = acc[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = threadIdx.x + j * kChanThreads;
            lds[(e >> log2M) * slots + lds_pad(e & (M - 1))] = acc[j];
        }
    }
    __syncthreads();
    const long long remaining = frames - t0;
    const int count = remaining < F ? static_cast<int>(remaining) : F;
    const int per = M >> 3;
    const int f = threadIdx.x / per;
    const int t = threadIdx.x - f * per;
    if (LOG2M == 6) {
        // The frame's two radix-8 passes (lds_pass's arithmetic) from registers: lane t of
        // frame f takes positions t + 8 r, runs the first pass, swaps through the frame's LDS
        // once and runs the second, ending with outputs t + 8 b (stockham_column's order for a
        // radix-8 last pass), which it stores itself: one LDS exchange instead of a pass's read
        // and write each plus the final read.
        const bool active = f < count;
        float2* fr = lds + f * slots;
        float2 u[8];
        if (active) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                u[r] = fr[lds_pad(t + 8 * r)];
            }
            stockham_pass_regs<true, 3>(u, t, 8, 1, table);
        }
        __syncthreads();  // every lane has read the branch sums it transforms
        stockham_exchange<3, 3, false>(fr, u, t, 8, 64, 1, active);
        if (active) {
            stockham_pass_regs<true, 3>(u, t, 8, 8, table);
            float2* zf = out + (t0 + f) * 64;
            // Non-temporal stores: the outputs are never re-read by the kernel, and kept out of
            // the way they leave L2 to the sample windows the neighbouring workgroups re-read
            // (2^24 samples, M = 64, 1024 taps: 61.7 -> 59.0 us per call).
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                __builtin_nontemporal_store(__builtin_bit_cast(double, u[b]),
                                            (global_double*)zf + t + 8 * b);  // address-space cast
            }
        }
        return;
    }
    lds_fft<true>(lds + f * slots, log2M, t, f < count, table);
    const int total = count * M;
    float2* dst = out + t0 * M;
    if (total == kChanOutputs) {
        // A full workgroup: the LDS reads of all eight rounds in flight together.
        constexpr int kRounds = kChanOutputs / kChanThreads;
        float2 z[kRounds];
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            const int e = threadIdx.x + k * kChanThreads;
            z[k] = lds[(e >> log2M) * slots + lds_pad(e & (M - 1))];
        }
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            gstore(dst, threadIdx.x + k * kChanThreads, z[k]);
        }
        return;
    }
    for (int e = threadIdx.x; e < total; e += kChanThreads) {
        dst[e] = lds[(e >> log2M) * slots + lds_pad(e & (M - 1))];
    }
}

std::size_t fused_lds_bytes(int log2M, int pb)
{
    const int M = 1 << log2M;
    const int F = kChanOutputs >> log2M;
    const std::size_t stage = static_cast<std::size_t>(F + pb - 1) * M + 2;  // + spare slots
    const std::size_t frames = static_cast<std::size_t>(F) * chan_frame_slots(log2M);
    return (stage > frames ? stage : frames) * sizeof(float2);
}

template <int PB>
void launch_fused(const float2* x, const float* taps, float2* out, int n, int log2M,
                  int num_taps, long long frames, std::size_t bytes, hipStream_t stream)
{
    const int F = kChanOutputs >> log2M;
    const unsigned grid = static_cast<unsigned>(8 * ceil_div(ceil_div(frames, F), 8));
    auto kernel = log2M == 6 ? channelize_fused_kernel<PB, 6> : channelize_fused_kernel<PB, 0>;
    if (bytes > 64 * 1024
        && hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(bytes)) != hipSuccess) {
        (void)hipGetLastError();  // not needed by this runtime; the launch reports problems
    }
    kernel<<<grid, kChanThreads, bytes, stream>>>(x, taps, out, n, log2M, num_taps, frames,
                                                  twiddle_table_4096());
    check_launch("channelize_fused_kernel");
}

// Dispatch on taps per branch: a compile-time register window for 1..16, generic otherwise.
bool try_fused(const float2* x, const float* taps, float2* out, int n, int M, int num_taps,
               long long frames, hipStream_t stream)
{
    if (!is_pow2(M) || M < 8 || M > kMaxLdsFft) {
        return false;
    }
    const int log2M = ilog2(M);
    const int pb = (num_taps + M - 1) / M;
    const std::size_t bytes = fused_lds_bytes(log2M, pb);
    if (bytes > static_cast<std::size_t>(max_lds_per_block())) {
        return false;
    }
    const bool windowed = pb <= kChanMaxWindowTaps && M <= kChanThreads;
    if (!windowed) {
        launch_fused<0>(x, taps, out, n, log2M, num_taps, frames, bytes, stream);
        return true;
    }
    switch (pb) {
#define RF_CHAN_CASE(P)                                                                        \
    case P:                                                                                    \
        launch_fused<P>(x, taps, out, n, log2M, num_taps, frames, bytes, stream);              \
        break;
        RF_CHAN_CASE(1) RF_CHAN_CASE(2) RF_CHAN_CASE(3) RF_CHAN_CASE(4)
        RF_CHAN_CASE(5) RF_CHAN_CASE(6) RF_CHAN_CASE(7) RF_CHAN_CASE(8)
        RF_CHAN_CASE(9) RF_CHAN_CASE(10) RF_CHAN_CASE(11) RF_CHAN_CASE(12)
        RF_CHAN_CASE(13) RF_CHAN_CASE(14) RF_CHAN_CASE(15) RF_CHAN_CASE(16)
#undef RF_CHAN_CASE
    default:
        launch_fused<0>(x, taps, out, n, log2M, num_taps, frames, bytes, stream);
    }
    return true;
}

}  // namespace
}  // namespace rf

using rf::guarded;

extern "C" {

int rf_channelize_frames(int num_samples, int channels)
{
    if (channels < 1 || num_samples < 0) {
        return 0;
    }
    return static_cast<int>((static_cast<long long>(num_samples) + channels - 1) / channels);
}

rf_status rf_channelize(const rf_fft_plan* plan, const rf_cf32* x, const float* taps,
                        rf_cf32* out, int num_samples, int channels, int num_taps,
                        rf_stream stream)
{
    return guarded([&] {
        if (channels < 1) rf::fail(RF_INVALID_ARG, "channelize: need at least one channel");
        if (num_taps < 1) rf::fail(RF_INVALID_ARG, "channelize: need at least one tap");
        if (num_samples < 0) rf::fail(RF_INVALID_ARG, "channelize: negative sample count");
        const long long frames = rf_channelize_frames(num_samples, channels);
        if (frames == 0) {
            return;
        }
        if (x == nullptr || taps == nullptr || out == nullptr) {
            rf::fail(RF_INVALID_ARG, "channelize: NULL buffer");
        }
        if (plan != nullptr && (plan->nfft != channels || plan->batch != frames)) {
            rf::fail(RF_INVALID_ARG, "channelize: the plan must be channels points x frames");
        }
        hipStream_t s = rf::as_stream(stream);
        const float2* fx = reinterpret_cast<const float2*>(x);
        float2* fo = reinterpret_cast<float2*>(out);
        if (rf::try_fused(fx, taps, fo, num_samples, channels, num_taps, frames, s)) {
            return;
        }
        const long long total = frames * channels;
        rf::channelize_branch_kernel<<<static_cast<unsigned>(rf::ceil_div(total, 256)), 256, 0,
                                       s>>>(fx, taps, fo, channels, num_taps, frames);
        rf::check_launch("channelize_branch_kernel");
        if (plan != nullptr) {
            rf::fft_exec(*plan, fo, fo, RF_FFT_INVERSE, s);
            return;
        }
        rf_fft_plan* own = nullptr;
        if (rf_fft_plan_create(channels, static_cast<int>(frames), &own) != RF_OK) {
            rf::fail(RF_ALLOC, std::string("channelize: ") + rf_last_error());
        }
        try {
            rf::fft_exec(*own, fo, fo, RF_FFT_INVERSE, s);
        } catch (...) {
            rf_fft_plan_destroy(own);
            throw;
        }
        // The plan's buffers must outlive the enqueued transform.
        RF_HIP_CHECK(hipStreamSynchronize(s));
        rf_fft_plan_destroy(own);
    });
}

}  // extern "C"
