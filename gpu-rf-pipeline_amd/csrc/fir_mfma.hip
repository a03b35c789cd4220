// Split-precision FIR on the matrix cores: a Toeplitz GEMM on v_mfma_f32_32x32x16_{f16,bf16}
// with a hi/lo split of both operands and FP32 accumulation.
//
// GEMM form of y[n] = sum_k h[k] x[n - k]: a wave owns 1024 consecutive outputs
// y[n_w + 32 m + c] (m, c < 32). With kk = k + c' shifted so that every column shares one
// A matrix,
//   Y[m][c] = sum_kk A[m][kk] B[kk][c],  A[m][kk] = x[n_w + 32 m + 31 - kk],
//                                        B[kk][c] = h[kk - 31 + c]  (0 outside 0..T-1),
// kk < Kp = ceil((T + 31) / 16) * 16, i.e. (T + 31) / T ~ 3 % more products than the direct
// sum. Each operand is split v = hi + lo in the 16-bit type and three products are
// accumulated (hi*hi + hi*lo + lo*hi); the dropped lo*lo term is 2^-22 of a product for
// FP16 and 2^-16 for BF16.
//
// FP16 range: the taps are scaled by a power of two so that max|h| is in [2^13, 2^14), and each
// workgroup scales its samples the same way from its own max|x|; the FP32 result is scaled
// back exactly. BF16 has FP32's exponent range and needs no scaling.
//
// LDS: the sample tile as four 16-bit arrays (re/im x hi/lo), so a lane's eight consecutive
// samples are one 16-byte read. The 32 lanes of an A read start 32 samples (64 bytes) apart,
// which would put every fourth lane on the same banks, so within each block of 32 samples the
// four groups of 8 are rotated by (block / 4) mod 4: lanes i .. i + 15 then read 16 different
// 16-byte bank groups, and no space is added. The taps are reversed and stored as eight copies shifted by one element each (hi and
// lo), so a lane's eight taps, whose start is not a multiple of eight, are still one aligned
// 16-byte read from the copy with the matching shift.

#include "fir_internal.hpp"

#include <string>

namespace rf {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaThreads = 256;                // 4 waves, 1024 outputs each
constexpr int kMfmaOutputs = 4 * 1024;           // per workgroup
constexpr int kMfmaMaxTaps = 2049;
// Per-lane load counts of the prologue at the largest supported tap count: tile samples
// (4096 + Kp - 32 <= 6144) and 16-byte tap-table vectors (2 copy_len <= 2 * 2136).
constexpr int kTileLoads = 24;
constexpr int kTapLoads = 17;

__host__ __device__ constexpr int mfma_kp(int num_taps)
{
    return ((num_taps + 31 + 15) / 16) * 16;
}

// LDS slot of tile sample u = 32 b + 8 g + e: group g moves to (g + b / 4) mod 4 within block
// b. A bijection on every block of 32; an 8-aligned group of 8 samples stays contiguous and
// 16-byte aligned.
__host__ __device__ constexpr int tile_slot(int u)
{
    return (u & ~31) | ((((u >> 3) + (u >> 7)) & 3) << 3) | (u & 7);
}

__host__ __device__ constexpr int tile_slots(int samples)
{
    return (samples + 31) & ~31;
}

// Length of one shifted tap copy: indices up to T + 78 are read.
__host__ __device__ constexpr int mfma_copy_len(int num_taps)
{
    return ((num_taps + 80 + 7) / 8) * 8;
}

static_assert(kMfmaOutputs + mfma_kp(kMfmaMaxTaps) - 32 <= kTileLoads * kMfmaThreads,
              "tile samples per lane");
static_assert(2 * mfma_copy_len(kMfmaMaxTaps) <= kTapLoads * kMfmaThreads,
              "tap-table vectors per lane");

template <bool BF16>
struct Split;

template <>
struct Split<false> {
    typedef _Float16 T;
    typedef f16x8 V;
    __device__ static void split(float v, T& hi, T& lo)
    {
        hi = static_cast<_Float16>(v);
        lo = static_cast<_Float16>(v - static_cast<float>(hi));
    }
    __device__ static f32x16 mma(V a, V b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Split<true> {
    typedef __bf16 T;
    typedef bf16x8 V;
    __device__ static void split(float v, T& hi, T& lo)
    {
        hi = static_cast<__bf16>(v);
        lo = static_cast<__bf16>(v - static_cast<float>(hi));
    }
    __device__ static f32x16 mma(V a, V b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

// Power-of-two exponent e with max * 2^e in [2^13, 2^14) (0 for max == 0).
__device__ __forceinline__ int fp16_scale_exponent(float max_abs)
{
    if (!(max_abs > 0.0f)) {
        return 0;
    }
    int e = 0;
    (void)frexpf(max_abs, &e);  // max = f 2^e, f in [0.5, 1)
    int s = 14 - e;
    return s > 100 ? 100 : (s < -100 ? -100 : s);
}

__device__ float block_max(float v, float* scratch)
{
    for (int off = 32; off > 0; off >>= 1) {
        v = fmaxf(v, __shfl_xor(v, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        scratch[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    float m = scratch[0];
    for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
        m = fmaxf(m, scratch[w]);
    }
    __syncthreads();
    return m;
}

// Workspace written by the prep kernel: hi copies, lo copies (8 x copy_len each), then one
// float with 2^-tap_exponent.
template <bool BF16>
__global__ __launch_bounds__(kMfmaThreads) void fir_mfma_taps_kernel(
    const float* __restrict__ taps, int num_taps, typename Split<BF16>::T* ws, float* descale)
{
    __shared__ float scratch[kMfmaThreads / 64];
    float m = 0.0f;
    for (int i = threadIdx.x; i < num_taps; i += blockDim.x) {
        m = fmaxf(m, fabsf(taps[i]));
    }
    m = block_max(m, scratch);
    const int e = BF16 ? 0 : fp16_scale_exponent(m);
    const int len = mfma_copy_len(num_taps);
    typename Split<BF16>::T* hi = ws;
    typename Split<BF16>::T* lo = ws + 8 * len;
    for (int idx = threadIdx.x; idx < 8 * len; idx += blockDim.x) {
        const int r = idx / len;
        const int v = idx - r * len;
        const int k = num_taps + 47 - v - r;  // copy[r][v] = h[T + 47 - v - r]
        const float h = (k >= 0 && k < num_taps) ? ldexpf(taps[k], e) : 0.0f;
        Split<BF16>::split(h, hi[idx], lo[idx]);
    }
    if (threadIdx.x == 0) {
        *descale = ldexpf(1.0f, -e);
    }
}

template <bool BF16>
__global__ __launch_bounds__(kMfmaThreads) void fir_mfma_kernel(
    const float2* __restrict__ x, const typename Split<BF16>::T* __restrict__ ws,
    const float* __restrict__ tap_descale, float2* __restrict__ y, int n, int num_taps)
{
    typedef typename Split<BF16>::T H;
    typedef typename Split<BF16>::V V;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ float scratch[kMfmaThreads / 64];
    const int kp = mfma_kp(num_taps);
    const int U = kMfmaOutputs + kp - 32;
    const int Us = tile_slots(U);
    const int len = mfma_copy_len(num_taps);
    H* rh = reinterpret_cast<H*>(smem);
    H* rl = rh + Us;
    H* ih = rl + Us;
    H* il = ih + Us;
    H* th = il + Us;          // 8 copies
    H* tl = th + 8 * len;     // 8 copies
    const long long n0 = static_cast<long long>(blockIdx.x) * kMfmaOutputs;
    const long long x0 = n0 - kp + 32;

    // Samples and taps go to LDS in chunks of kChunk loads per lane, so that the prologue holds
    // at most kChunk values in registers: with all of a lane's loads in flight at once it took
    // more registers than the matrix-core loop (228 VGPRs and 64 AGPRs, one wave per SIMD).
    // FP16 first needs the block's max |x| for its scale: one pass over the samples for the
    // max, a second (L2 hits) to split them; BF16 splits in the first pass.
    constexpr int kChunk = 8;
    static_assert(kTileLoads % kChunk == 0, "whole chunks of tile loads");
    auto sample = [&](int u) {
        const long long s = x0 + u;
        return (u < U && s >= 0 && s < n) ? x[s] : make_float2(0.0f, 0.0f);
    };
    auto put = [&](int u, float2 v, int ex) {
        if (u < U) {
            const int slot = tile_slot(u);
            Split<BF16>::split(ldexpf(v.x, ex), rh[slot], rl[slot]);
            Split<BF16>::split(ldexpf(v.y, ex), ih[slot], il[slot]);
        }
    };
    int ex = 0;
    if (!BF16) {
        float m = 0.0f;
#pragma unroll 1
        for (int c = 0; c < kTileLoads; c += kChunk) {
            float2 xs[kChunk];
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                xs[j] = sample(threadIdx.x + (c + j) * kMfmaThreads);
            }
#pragma unroll
            for (int j = 0; j < kChunk; ++j) {
                m = fmaxf(m, fmaxf(fabsf(xs[j].x), fabsf(xs[j].y)));
            }
        }
        ex = fp16_scale_exponent(block_max(m, scratch));
    }
#pragma unroll 1
    for (int c = 0; c < kTileLoads; c += kChunk) {
        float2 xs[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            xs[j] = sample(threadIdx.x + (c + j) * kMfmaThreads);
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            put(threadIdx.x + (c + j) * kMfmaThreads, xs[j], ex);
        }
    }
    // Taps: 16-byte copies of the hi/lo tables (16 len 16-bit values); th and tl are
    // contiguous, as in the workspace.
    const int vecs = 2 * len;  // 16 len values / 8 per 16-byte vector
    const u32x4* wsv = reinterpret_cast<const u32x4*>(ws);
    u32x4* thv = reinterpret_cast<u32x4*>(th);
#pragma unroll 1
    for (int c = 0; c < kTapLoads; c += kChunk) {
        u32x4 tv[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int idx = threadIdx.x + (c + j) * kMfmaThreads;
            if (c + j < kTapLoads && idx < vecs) {
                tv[j] = wsv[idx];
            }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int idx = threadIdx.x + (c + j) * kMfmaThreads;
            if (c + j < kTapLoads && idx < vecs) {
                thv[idx] = tv[j];
            }
        }
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int i = lane & 31;
    const int g = lane >> 5;
    f32x16 acc_re0 = {};
    f32x16 acc_im0 = {};
    f32x16 acc_re1 = {};
    f32x16 acc_im1 = {};
    // Step s covers kk = 16 s .. 16 s + 15; even steps accumulate into acc*0, odd ones into
    // acc*1, so the accumulators stay in place. A step's six fragments (sample hi/lo of re and
    // im, tap hi/lo) are read from LDS one step ahead: step s + 1's while step s multiplies,
    // so the LDS latency hides behind six matrix-core instructions instead of stalling each
    // step (with the reads next to their MFMAs the loop alone ran at 43 % of the matrix-core
    // rate, using under half the LDS bandwidth).
    struct Frags {
        V arh, arl, aih, ail, bh, bl;
    };
    auto load = [&](int s) {
        Frags f;
        const int u0 = tile_slot(1024 * w + 32 * i - 16 * s - 8 * g + kp - 8);
        f.arh = *reinterpret_cast<const V*>(rh + u0);
        f.arl = *reinterpret_cast<const V*>(rl + u0);
        f.aih = *reinterpret_cast<const V*>(ih + u0);
        f.ail = *reinterpret_cast<const V*>(il + u0);
        const int b = num_taps + 71 - 16 * s - 8 * g - i;  // b0 + 48 > 0
        const int r = b & 7;
        const int base = r * len + (b - r);
        f.bh = *reinterpret_cast<const V*>(th + base);
        f.bl = *reinterpret_cast<const V*>(tl + base);
        return f;
    };
    auto mma = [](const Frags& f, f32x16& acc_re, f32x16& acc_im) {
        acc_re = Split<BF16>::mma(f.arh, f.bh, acc_re);
        acc_im = Split<BF16>::mma(f.aih, f.bh, acc_im);
        acc_re = Split<BF16>::mma(f.arh, f.bl, acc_re);
        acc_im = Split<BF16>::mma(f.aih, f.bl, acc_im);
        acc_re = Split<BF16>::mma(f.arl, f.bh, acc_re);
        acc_im = Split<BF16>::mma(f.ail, f.bh, acc_im);
    };
    const int steps = kp / 16;
    int s = 0;
    Frags f0 = load(0);
    for (; s + 1 < steps; s += 2) {
        const Frags f1 = load(s + 1);
        mma(f0, acc_re0, acc_im0);
        if (s + 2 < steps) {
            f0 = load(s + 2);
        }
        mma(f1, acc_re1, acc_im1);
    }
    if (s < steps) {
        mma(f0, acc_re0, acc_im0);
    }
    const float scale = ldexpf(*tap_descale, -ex);
    const long long nw = n0 + 1024 * w;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = (v & 3) + 8 * (v >> 2) + 4 * g;
        const long long o = nw + 32 * row + i;
        if (o < n) {
            y[o] = make_float2((acc_re0[v] + acc_re1[v]) * scale, (acc_im0[v] + acc_im1[v]) * scale);
        }
    }
}

std::size_t mfma_lds_bytes(int num_taps)
{
    const std::size_t Us = tile_slots(kMfmaOutputs + mfma_kp(num_taps) - 32);
    return (4 * Us + 16 * static_cast<std::size_t>(mfma_copy_len(num_taps))) * 2;
}

template <bool BF16>
void launch_mfma(const float2* x, const float* taps, float2* y, int n, int num_taps,
                 hipStream_t stream, void* workspace)
{
    if (num_taps > kMfmaMaxTaps) {
        fail(RF_LIMIT, "fir mfma: at most " + std::to_string(kMfmaMaxTaps) + " taps");
    }
    const std::size_t bytes = mfma_lds_bytes(num_taps);
    if (bytes > static_cast<std::size_t>(max_lds_per_block())) {
        fail(RF_LIMIT, "fir mfma: " + std::to_string(num_taps) + " taps need "
                           + std::to_string(bytes) + " bytes of LDS per workgroup");
    }
    const std::size_t ws_bytes = fir_mfma_workspace_bytes(num_taps);
    if (ws_bytes > kStreamWorkspaceBytes) {
        fail(RF_LIMIT, "fir mfma: " + std::to_string(num_taps) + " taps need "
                           + std::to_string(ws_bytes) + " bytes of workspace");
    }
    typedef typename Split<BF16>::T H;
    H* tabs = static_cast<H*>(workspace);
    float* descale = reinterpret_cast<float*>(static_cast<char*>(workspace) + ws_bytes - 16);
    fir_mfma_taps_kernel<BF16><<<1, kMfmaThreads, 0, stream>>>(taps, num_taps, tabs, descale);
    check_launch("fir_mfma_taps_kernel");
    allow_large_lds(reinterpret_cast<const void*>(fir_mfma_kernel<BF16>), bytes);
    const unsigned grid = static_cast<unsigned>(ceil_div(n, kMfmaOutputs));
    fir_mfma_kernel<BF16><<<grid, kMfmaThreads, bytes, stream>>>(x, tabs, descale, y, n,
                                                                  num_taps);
    check_launch("fir_mfma_kernel");
}

}  // namespace

std::size_t fir_mfma_workspace_bytes(int num_taps)
{
    // 16 split tap copies (hi, lo) of mfma_copy_len halves each, then 16 bytes of descale
    return 16 * static_cast<std::size_t>(mfma_copy_len(num_taps)) * 2 + 16;
}

void fir_mfma_launch(bool bf16, const float2* x, const float* taps, float2* y, int n,
                     int num_taps, hipStream_t stream, void* workspace)
{
    if (bf16) {
        launch_mfma<true>(x, taps, y, n, num_taps, stream, workspace);
    } else {
        launch_mfma<false>(x, taps, y, n, num_taps, stream, workspace);
    }
}

}  // namespace rf
