// FIR filter kernels for gfx950 and the rf_fir / rf_fir_batched entry points.
//
// Definition (src/fir.cu:1-5, tools/reference.py:15-17,103-111):
//   y[n] = sum_{k=0}^{T-1} taps[k] * x[n - k],  x[m] = 0 for m < 0,  output length N.
//
// The direct variants (naive, tiled, tiled_const, register) all accumulate
//   acc = fma(taps[k], x[n-k], acc)  for k = 0, 1, ..., T-1, starting from 0
// per output and per component, so they give identical bits (up to the sign of a zero). The
// reference's variant table is kept (src/fir.cu:725-734) with its launch semantics; the
// kernels themselves are laid out for 64-lane wavefronts:
//   - tiles are filled with 16-byte loads (two samples per lane),
//   - tap reads are wave-uniform and go through the scalar cache (s_load), which is what the
//     reference's __constant__ bank is for, without any program-wide tap state,
//   - the register variant keeps R = 4, 8 or 16 outputs per lane and walks the taps R at a
//     time, reading one 8R-byte window per R taps from an LDS tile padded by 16 bytes per
//     R samples (conflict-free 16-byte LDS reads).
// fft_ola is new: overlap-save with a 4096-point transform for long filters and a 512-point
// one held by a single wavefront for short ones.

#include "fft4096.hpp"
#include "fft512.hpp"
#include "fft_core.hpp"
#include "fir_internal.hpp"
#include "rf_common.hpp"
#include "twiddle.hpp"

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace rf {
namespace {

// Pointer into the constant address space: wave-uniform loads through it become s_load
// (scalar data cache), the AMD counterpart of the reference's constant-memory taps.
typedef const __attribute__((address_space(4))) float* const_taps_ptr;

__device__ __forceinline__ const_taps_ptr as_const_taps(const float* p)
{
    return (const_taps_ptr)(p);
}

__device__ __forceinline__ float2 load_sample(const float2* __restrict__ x, long long m, int n)
{
    return (m >= 0 && m < n) ? x[m] : make_float2(0.0f, 0.0f);
}

// --- naive ------------------------------------------------------------------------------

// One lane per output; taps and samples from global memory (src/fir.cu:20-43).
__device__ __forceinline__ void fir_naive_one(const float2* __restrict__ x,
                                              const float* __restrict__ taps,
                                              float2* __restrict__ y, int n, int num_taps,
                                              long long i)
{
    if (i >= n) {
        return;
    }
    const int last = i < num_taps - 1 ? static_cast<int>(i) : num_taps - 1;
    float re = 0.0f;
    float im = 0.0f;
    for (int k = 0; k <= last; ++k) {
        const float h = taps[k];
        const float2 s = x[i - k];
        re = fmaf(h, s.x, re);
        im = fmaf(h, s.y, im);
    }
    y[i] = make_float2(re, im);
}

__global__ void fir_naive_kernel(const float2* __restrict__ x, const float* __restrict__ taps,
                                 float2* __restrict__ y, int n, int num_taps)
{
    fir_naive_one(x, taps, y, n, num_taps,
                  static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

// --- tiled / tiled_const ------------------------------------------------------------------

// Tile of halo + blockDim.x samples: tile[i] = x[block_start - halo + i], zero outside the
// input (src/fir.cu:89-99). Filled two samples per lane per step.
__device__ __forceinline__ void fill_tile(const float2* __restrict__ x, int n, long long start,
                                          int len, float2* tile)
{
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        tile[i] = load_sample(x, start + i, n);
    }
}

template <bool CONST_TAPS>
__device__ __forceinline__ void fir_tiled_block(const float2* __restrict__ x,
                                                const float* __restrict__ taps,
                                                float2* __restrict__ y, int n, int num_taps,
                                                float2* tile)
{
    const int halo = num_taps - 1;
    const long long block_start = static_cast<long long>(blockIdx.x) * blockDim.x;
    fill_tile(x, n, block_start - halo, halo + static_cast<int>(blockDim.x), tile);
    __syncthreads();
    const long long i = block_start + threadIdx.x;
    if (i >= n) {
        return;
    }
    const int newest = halo + threadIdx.x;
    float re = 0.0f;
    float im = 0.0f;
    if (CONST_TAPS) {
        const const_taps_ptr ct = as_const_taps(taps);
        for (int k = 0; k < num_taps; ++k) {
            const float h = ct[k];
            const float2 s = tile[newest - k];
            re = fmaf(h, s.x, re);
            im = fmaf(h, s.y, im);
        }
    } else {
        for (int k = 0; k < num_taps; ++k) {
            const float h = taps[k];
            const float2 s = tile[newest - k];
            re = fmaf(h, s.x, re);
            im = fmaf(h, s.y, im);
        }
    }
    y[i] = make_float2(re, im);
}

template <bool CONST_TAPS>
__global__ void fir_tiled_kernel(const float2* __restrict__ x, const float* __restrict__ taps,
                                 float2* __restrict__ y, int n, int num_taps)
{
    extern __shared__ float2 tile[];
    fir_tiled_block<CONST_TAPS>(x, taps, y, n, num_taps, tile);
}

// --- register -----------------------------------------------------------------------------

// LDS slot of tile sample s: groups of R samples, each followed by two pad slots (16 bytes),
// so group g starts at g * (R + 2) and stays 16-byte aligned.
template <int R>
__device__ __forceinline__ int reg_slot(int s)
{
    return (s / R) * (R + 2) + (s % R);
}

// R outputs per lane. The tile holds x[block_start - Tpad .. block_start + blockDim.x * R)
// with Tpad = ceil(T / R) * R, so lane l's window for taps [cR, cR + R) is group
// (Tpad/R + l - c) ("cur") and the group before it ("prev"):
//   x[first + j - k] = (j >= i) ? cur[j - i] : prev[R + j - i],   k = cR + i.
// `block` is the index of the blockDim.x * R outputs this workgroup computes.
template <int R>
__device__ __forceinline__ void fir_register_block(const float2* __restrict__ x,
                                                   const float* __restrict__ taps,
                                                   float2* __restrict__ y, int n, int num_taps,
                                                   float2* tile, long long block)
{
    const int outputs = static_cast<int>(blockDim.x) * R;
    const int chunks = (num_taps + R - 1) / R;
    const int tpad = chunks * R;
    const int len = tpad + outputs;  // a multiple of R, so even
    const long long block_start = block * outputs;
    const long long tile_start = block_start - tpad;

    for (int s = 2 * threadIdx.x; s < len; s += 2 * blockDim.x) {
        const long long m = tile_start + s;
        float4 v;
        if (m >= 0 && m + 1 < n) {
            const float2 a = x[m];
            const float2 b = x[m + 1];
            v = make_float4(a.x, a.y, b.x, b.y);
        } else {
            const float2 a = load_sample(x, m, n);
            const float2 b = load_sample(x, m + 1, n);
            v = make_float4(a.x, a.y, b.x, b.y);
        }
        *reinterpret_cast<float4*>(&tile[reg_slot<R>(s)]) = v;
    }
    __syncthreads();

    const long long first = block_start + static_cast<long long>(threadIdx.x) * R;
    if (first >= n) {
        return;
    }
    const const_taps_ptr ct = as_const_taps(taps);
    float2 cur[R];
    float2 prev[R];
    float2 acc[R];
    {
        const float4* g = reinterpret_cast<const float4*>(tile + (chunks + threadIdx.x) * (R + 2));
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            const float4 v = g[q];
            cur[2 * q] = make_float2(v.x, v.y);
            cur[2 * q + 1] = make_float2(v.z, v.w);
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        acc[j] = make_float2(0.0f, 0.0f);
    }
    const int full = num_taps / R;  // chunks whose R taps all exist
    for (int c = 0; c < chunks; ++c) {
        {
            const float4* g =
                reinterpret_cast<const float4*>(tile + (chunks + threadIdx.x - c - 1) * (R + 2));
#pragma unroll
            for (int q = 0; q < R / 2; ++q) {
                const float4 v = g[q];
                prev[2 * q] = make_float2(v.x, v.y);
                prev[2 * q + 1] = make_float2(v.z, v.w);
            }
        }
        float h[R];
        if (c < full) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                h[i] = ct[c * R + i];
            }
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int k = c * R + i;
                h[i] = k < num_taps ? ct[k] : 0.0f;
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const float2 v = (j >= i) ? cur[j - i] : prev[R + j - i];
                acc[j].x = fmaf(h[i], v.x, acc[j].x);
                acc[j].y = fmaf(h[i], v.y, acc[j].y);
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            cur[j] = prev[j];
        }
    }
    const bool aligned = (reinterpret_cast<uintptr_t>(y) & 15) == 0;
    if (aligned && first + R <= n) {
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            *reinterpret_cast<float4*>(&y[first + 2 * q]) =
                make_float4(acc[2 * q].x, acc[2 * q].y, acc[2 * q + 1].x, acc[2 * q + 1].y);
        }
    } else {
#pragma unroll
        for (int j = 0; j < R; ++j) {
            if (first + j < n) {
                y[first + j] = acc[j];
            }
        }
    }
}

template <int R>
__global__ void fir_register_kernel(const float2* __restrict__ x, const float* __restrict__ taps,
                                    float2* __restrict__ y, int n, int num_taps)
{
    extern __shared__ float2 tile[];
    fir_register_block<R>(x, taps, y, n, num_taps, tile, blockIdx.x);
}

// --- batched launches (src/fir_batch.hpp:9-13, src/fused.cu:36-74, src/fir.cu:562-721) ----

struct FirBatch {
    const float2* x[RF_MAX_FIR_BATCH];
    float2* y[RF_MAX_FIR_BATCH];
    int n[RF_MAX_FIR_BATCH];
};

__global__ void fir_naive_batched_kernel(FirBatch batch, const float* __restrict__ taps,
                                         int num_taps)
{
    const int c = blockIdx.y;
    fir_naive_one(batch.x[c], taps, batch.y[c], batch.n[c], num_taps,
                  static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x);
}

template <bool CONST_TAPS>
__global__ void fir_tiled_batched_kernel(FirBatch batch, const float* __restrict__ taps,
                                         int num_taps)
{
    extern __shared__ float2 tile[];
    const int c = blockIdx.y;
    // A block wholly past its channel's end returns before the barrier; every lane of it
    // does, so the barrier is never split.
    if (static_cast<long long>(blockIdx.x) * blockDim.x >= batch.n[c]) {
        return;
    }
    fir_tiled_block<CONST_TAPS>(batch.x[c], taps, batch.y[c], batch.n[c], num_taps, tile);
}

template <int R>
__global__ void fir_register_batched_kernel(FirBatch batch, const float* __restrict__ taps,
                                            int num_taps)
{
    extern __shared__ float2 tile[];
    const int c = blockIdx.y;
    if (static_cast<long long>(blockIdx.x) * blockDim.x * R >= batch.n[c]) {
        return;
    }
    fir_register_block<R>(batch.x[c], taps, batch.y[c], batch.n[c], num_taps, tile, blockIdx.x);
}

// --- fft_ola --------------------------------------------------------------------------------

// Overlap-save: with P = 4096 and Q = P - (T - 1) outputs per block,
//   seg[i] = x[bQ - (T-1) + i],  i < P (zero outside the input)
//   y[bQ + j] = IFFT(FFT(seg) * FFT(h)) [T - 1 + j] / P,  j < Q.
// 256 lanes per segment with the register-resident radix-16 transform of fft4096.hpp: the
// forward transform reads x straight from HBM, the spectrum multiply sits between the forward
// and the inverse transform in registers, and the inverse transform writes y straight back.
// The tap spectrum is made once per call by one workgroup with the same transform.
constexpr int kOlaP = k4096;
constexpr int kOlaMinTaps = 65;   // the matrix-core variants start above this
constexpr int kOlaMaxTaps = 2049; // keeps Q >= P / 2

// Tap spectrum, scaled by 1/P (exact: a power of two): one workgroup, the register-resident
// radix-16 transform.
// Tap spectrum H[k] = DFT_4096(h / 4096) (the inverse transform's scale folded in), bins paired
// as the overlap-save kernel reads them: bins t + 256 r and t + 256 (r + 1), r even, at
// spec + 2 ((r / 2) 256 + t), so that one 16-byte load per lane, contiguous over lanes, gives both.
__device__ __forceinline__ void ola_taps_spectrum(const float* __restrict__ taps, int num_taps,
                                                 float2* __restrict__ spec, float2* buf,
                                                 const float2* __restrict__ stages)
{
    const int t = threadIdx.x;
    const float scale = 1.0f / static_cast<float>(kOlaP);
    // range-checked buffer loads: taps past num_taps read as zero, and all 16 loads are in
    // flight together (a per-lane test around each load would wait for each in turn)
    const __amdgpu_buffer_rsrc_t h = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(taps), static_cast<short>(0), num_taps * 4, 0x00020000);
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const float v = __builtin_bit_cast(
            float, __builtin_amdgcn_raw_buffer_load_b32(h, static_cast<unsigned>(t) * 4u,
                                                        static_cast<unsigned>(r * k4096Lanes) * 4u, 0));
        u[r] = make_float2(v * scale, 0.0f);
    }
    fft4096<false>(buf, t, u, stages);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        spec[2 * ((r >> 1) * k4096Lanes + t) + (r & 1)] = u[r];
    }
}

__global__ __launch_bounds__(k4096Lanes) void fir_ola_taps_kernel(
    const float* __restrict__ taps, int num_taps, float2* __restrict__ spec,
    const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    ola_taps_spectrum(taps, num_taps, spec, buf, stages);
}

// Unchecked buffer accesses for blocks known to lie inside their arrays: address = base +
// soffset + voffset, with the workgroup-uniform parts (base, block offset) in scalar registers
// and one per-lane offset register shared by all 16 blocks of a segment, instead of a 64-bit
// address per access. The record count is the maximum, so the range check never applies.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t unchecked_buffer(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0), -1,
                                             0x00020000);
}

__device__ __forceinline__ float2 buffer_load2(__amdgpu_buffer_rsrc_t r, unsigned voffset,
                                               unsigned soffset)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voffset, soffset, 0));
}

// Non-temporal (cache policy bit 1, "nt"): the outputs are streamed out and never re-read by
// the kernel, so they should not displace the input halos the next segments read from L2.
__device__ __forceinline__ void buffer_store2(float2 v, __amdgpu_buffer_rsrc_t r,
                                              unsigned voffset, unsigned soffset)
{
    using u32x2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voffset, soffset, 2);
}

// A lane's 16 samples x[in0 + t + 256 r] of a segment (zero outside the input).
__device__ __forceinline__ void ola_load(const float2* __restrict__ x, int n, long long in0, int t,
                                         float2 (&u)[16])
{
    const unsigned lane_bytes = static_cast<unsigned>(t) * 8u;
    constexpr unsigned kBlockBytes = k4096Lanes * 8u;
    if (in0 >= 0 && in0 + kOlaP <= n
        && (reinterpret_cast<unsigned long long>(x + in0) & 15ULL) == 0) {
        // Blocks 2 q and 2 q + 1 with one 16-byte load per lane: an even lane t reads samples
        // t, t + 1 of block 2 q, an odd lane samples t - 1, t of block 2 q + 1; one DPP swap
        // between the lane pair gives each lane sample t of both blocks.
        const __amdgpu_buffer_rsrc_t src = unchecked_buffer(x + in0);
        const bool odd = (t & 1) != 0;
        const unsigned pair_bytes = odd ? (lane_bytes - 8u + kBlockBytes) : lane_bytes;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            typedef float f32x4 __attribute__((ext_vector_type(4)));
            const f32x4 v = __builtin_bit_cast(
                f32x4, __builtin_amdgcn_raw_buffer_load_b128(src, pair_bytes, q * 2u * kBlockBytes, 0));
            const float sx = odd ? v.x : v.z;   // even lane sends sample t + 1 of block 2 q
            const float sy = odd ? v.y : v.w;   // odd lane sends sample t - 1 of block 2 q + 1
            const float rx = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(
                                                           __builtin_bit_cast(int, sx), 0xB1, 0xF, 0xF, false));
            const float ry = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(
                                                           __builtin_bit_cast(int, sy), 0xB1, 0xF, 0xF, false));
            u[2 * q] = odd ? make_float2(rx, ry) : make_float2(v.x, v.y);
            u[2 * q + 1] = odd ? make_float2(v.z, v.w) : make_float2(rx, ry);
        }
    } else if (in0 >= 0 && in0 + kOlaP <= n) {
        const __amdgpu_buffer_rsrc_t src = unchecked_buffer(x + in0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            u[r] = buffer_load2(src, lane_bytes, r * kBlockBytes);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            u[r] = load_sample(x, in0 + t + r * k4096Lanes, n);
        }
    }
}

// u (the forward transform of a segment) times the tap spectrum, read in bin pairs
// (ola_taps_spectrum): 8 loads of 16 bytes per lane.
__device__ __forceinline__ void ola_spectrum_product(const float2* __restrict__ spec, int t,
                                                     float2 (&u)[16])
{
    constexpr unsigned kBlockBytes = k4096Lanes * 8u;
    const __amdgpu_buffer_rsrc_t h = unchecked_buffer(spec);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4 hv = __builtin_bit_cast(
            f32x4, __builtin_amdgcn_raw_buffer_load_b128(h, 2u * static_cast<unsigned>(t) * 8u,
                                                         q * 2u * kBlockBytes, 0));
        u[2 * q] = cmul(u[2 * q], make_float2(hv.x, hv.y));
        u[2 * q + 1] = cmul(u[2 * q + 1], make_float2(hv.z, hv.w));
    }
}

// Outputs of a segment: position p = t + 256 r goes to y[out0 + p - halo] when p >= halo (and
// inside the output).
__device__ __forceinline__ void ola_store(float2* __restrict__ y, int n, int halo, long long out0,
                                          const float2 (&u)[16], int t)
{
    const unsigned lane_bytes = static_cast<unsigned>(t) * 8u;
    if (out0 + kOlaP - halo <= n) {
        // a block wholly at or above the halo stores without a per-lane test
        const __amdgpu_buffer_rsrc_t dst = unchecked_buffer(y + out0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int first = r * k4096Lanes;
            if (first >= halo) {
                buffer_store2(u[r], dst, lane_bytes, static_cast<unsigned>(first - halo) * 8u);
            } else if (first + k4096Lanes > halo && t + first >= halo) {
                buffer_store2(u[r], dst, static_cast<unsigned>(t + first - halo) * 8u, 0u);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int pos = t + r * k4096Lanes;
            const long long o = out0 + pos - halo;
            if (pos >= halo && o < n) {
                y[o] = u[r];
            }
        }
    }
}

// One segment: seg[i] = x[in0 + i], i < 4096 (zero outside the input), filtered by the
// spectrum, outputs y[out0 + i - (T-1)] for T-1 <= i < 4096. Lane t (threadIdx.x, passed in)
// holds positions t + 256 r. The twiddle factors are loaded after the samples, so they arrive
// while the first pass runs, and serve both transforms.
__device__ __forceinline__ void ola_segment(const float2* __restrict__ x,
                                            const float2* __restrict__ spec,
                                            float2* __restrict__ y, int n, int halo,
                                            long long out0, float2* buf,
                                            const float2* __restrict__ stages, int t)
{
    float2 u[16];
    ola_load(x, n, out0 - halo, t, u);
    Fft4096Twiddles tw;
    fft4096_load_twiddles(t, stages, tw);
    fft4096<false>(buf, t, u, tw);
    ola_spectrum_product(spec, t, u);
    __syncthreads();  // every lane has read the forward result before it is overwritten
    fft4096<true>(buf, t, u, tw);
    ola_store(y, n, halo, out0, u, t);
}

// Segment of a workgroup, XCD-major: workgroups are dispatched round-robin over the 8 XCDs, so
// workgroup b runs on XCD b % 8; giving XCD x the consecutive segments x*per .. x*per + per - 1
// in order makes each segment's (T-1)-sample halo, read by the previous segment, an L2 hit.
__device__ __forceinline__ long long xcd_major(long long b, long long per)
{
    return (b & 7) * per + (b >> 3);
}

__global__ __launch_bounds__(k4096Lanes) void fir_ola_kernel(const float2* __restrict__ x,
                                                           const float2* __restrict__ spec,
                                                           float2* __restrict__ y, int n,
                                                           int num_taps, long long segments,
                                                           long long per_xcd,
                                                           const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    const long long seg = xcd_major(blockIdx.x, per_xcd);
    if (seg >= segments) {
        return;  // the whole workgroup: padding of the grid to a multiple of 8
    }
    const int halo = num_taps - 1;
    ola_segment(x, spec, y, n, halo, seg * (kOlaP - halo), buf, stages, threadIdx.x);
}

__global__ __launch_bounds__(k4096Lanes) void fir_ola_batched_kernel(
    FirBatch batch, const float2* __restrict__ spec, int num_taps,
    const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    const int c = blockIdx.y;
    const int halo = num_taps - 1;
    const long long out0 = static_cast<long long>(blockIdx.x) * (kOlaP - halo);
    if (out0 >= batch.n[c]) {
        return;
    }
    ola_segment(batch.x[c], spec, batch.y[c], batch.n[c], halo, out0, buf, stages, threadIdx.x);
}

// --- persistent batch over any number of uneven captures --------------------------------------

// One capture of a persistent FIR batch.
struct FirCapture {
    const float2* x;
    float2* y;
    int n;
};

// Work items {capture, segment} (overlap-save) or {capture, output block} (register), built on
// the host in capture order, walked XCD-major: each XCD takes a contiguous run of items, so a
// segment's halo was just read by its predecessor behind the same L2.
// The lane index goes through an empty asm inside the walk: without it the compiler hoists
// every per-lane address of the segment out of the loop into held registers (166 VGPRs, three
// workgroups per CU); with it and a four-wave bound the kernel keeps the single-capture
// kernel's four workgroups per CU (111 VGPRs, no spills).
__global__ __launch_bounds__(k4096Lanes, 4) void fir_ola_batch_kernel(
    const FirCapture* __restrict__ caps, const int2* __restrict__ items, int total,
    const float2* __restrict__ spec, int num_taps, const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    const int halo = num_taps - 1;
    xcd_major_items(total, [&](int item) {
        const int2 it = items[item];
        const FirCapture c = caps[it.x];
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        ola_segment(c.x, spec, c.y, c.n, halo, static_cast<long long>(it.y) * (kOlaP - halo), buf,
                    stages, t);
        __syncthreads();  // the next segment reuses buf
    });
}

template <int R>
__global__ void fir_register_batch_kernel(const FirCapture* __restrict__ caps,
                                          const int2* __restrict__ items, int total,
                                          const float* __restrict__ taps, int num_taps)
{
    extern __shared__ float2 tile[];
    xcd_major_items(total, [&](int item) {
        const int2 it = items[item];
        const FirCapture c = caps[it.x];
        fir_register_block<R>(c.x, taps, c.y, c.n, num_taps, tile, it.y);
        __syncthreads();  // the next block refills the tile
    });
}

// --- fft_ola for short filters ----------------------------------------------------------------

// Overlap-save as above with a 512-point transform held by one wavefront (fft512.hpp), for
// kOla512MinTaps <= T <= kOla512MaxTaps: Q = 513 - T outputs per segment (288 at 225 taps).
// The direct register kernel costs 2 T FMA per output component; this form about 60 VALU
// instructions per output at 65 taps (two transforms of three radix-8 passes and one product
// per bin, over Q outputs), against about 110 for the 4096-point form, which overtakes it
// only where Q has shrunk to about half a segment (measured on 2^24 samples: 63.6 against
// 76.4 us at 66 taps, 62.4 against 66.7 at 193, 74.7 against 66.6 at 257). A workgroup is one
// wave (64 lanes), so a segment's barriers hold only its own wave. Each persistent wave makes
// the tap spectrum itself before its first segment (one transform more per wave, where a
// separate spectrum launch cost about 2.6 us per call on 2^20-sample captures), so every call
// is one launch and needs no device memory beyond its inputs and outputs; the wave then keeps
// its 14 transform factors and 8 spectrum bins in registers across its segments, and loads
// its next segment's samples before transforming the current one.
constexpr int kOla512P = k512;
constexpr int kOla512MinTaps = 17;
constexpr int kOla512MaxTaps = 225;
constexpr int kOla512WavesPerCu = 16;

// What a lane keeps across the segments it filters.
struct Ola512Lane {
    Fft512Twiddles tw;
    float2 h[8];  // tap spectrum H[t + 64 r], H = DFT_512(h / 512)
};

// The lane's factors and spectrum bins; ends with a barrier (the transform used buf).
__device__ __forceinline__ void ola512_lane_setup(const float* __restrict__ taps, int num_taps,
                                                  const float2* __restrict__ table, int t,
                                                  float2* buf, Ola512Lane& s)
{
    fft512_load_twiddles(t, table, s.tw);
    const float scale = 1.0f / static_cast<float>(kOla512P);  // exact: a power of two
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int k = t + r * k512Lanes;
        s.h[r] = make_float2(k < num_taps ? taps[k] * scale : 0.0f, 0.0f);
    }
    fft512<false>(buf, t, s.h, s.tw);
    lds_barrier();
}

// A segment: its capture and the index of its first output; seg[i] = x[out0 - halo + i],
// i < 512 (zero outside the input), gives y[out0 + i - halo] for halo <= i < 512.
struct Ola512Seg {
    const float2* x;
    float2* y;
    int n;
    long long out0;
};

__device__ __forceinline__ void ola512_load(const Ola512Seg& g, int halo, int t, float2 (&u)[8])
{
    const long long in0 = g.out0 - halo;
    if (in0 >= 0 && in0 + kOla512P <= g.n) {
        const __amdgpu_buffer_rsrc_t src = unchecked_buffer(g.x + in0);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            u[r] = buffer_load2(src, static_cast<unsigned>(t) * 8u,
                                static_cast<unsigned>(r * k512Lanes) * 8u);
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            u[r] = load_sample(g.x, in0 + t + r * k512Lanes, g.n);
        }
    }
}

// Filters the loaded segment u and stores its outputs; ends with a barrier before buf is
// written again.
__device__ __forceinline__ void ola512_filter(const Ola512Seg& g, int halo, float2 (&u)[8],
                                              float2* buf, const Ola512Lane& s, int t)
{
    fft512<false>(buf, t, u, s.tw);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        u[r] = cmul(u[r], s.h[r]);
    }
    lds_barrier();  // every lane has read the forward transform's last exchange
    fft512<true>(buf, t, u, s.tw);
    const unsigned lane_bytes = static_cast<unsigned>(t) * 8u;
    if (g.out0 + kOla512P - halo <= g.n) {
        // position p = t + 64 r is an output when p >= halo: blocks wholly above the halo store
        // without a per-lane test
        const __amdgpu_buffer_rsrc_t dst = unchecked_buffer(g.y + g.out0);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int first = r * k512Lanes;
            if (first >= halo) {
                buffer_store2(u[r], dst, lane_bytes, static_cast<unsigned>(first - halo) * 8u);
            } else if (first + k512Lanes > halo && t + first >= halo) {
                buffer_store2(u[r], dst, static_cast<unsigned>(t + first - halo) * 8u, 0u);
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int pos = t + r * k512Lanes;
            const long long o = g.out0 + pos - halo;
            if (pos >= halo && o < g.n) {
                g.y[o] = u[r];
            }
        }
    }
    lds_barrier();
}

// A persistent wave's walk over items 0 .. total - 1 (XCD-major, as xcd_major_items), with
// seg_of(item) giving each item's segment. The samples of the wave's next segment are loaded
// before the current one is transformed (and the first segment's before the tap spectrum), so
// the wave keeps a segment's loads in flight behind its transforms.
template <class SegOf>
__device__ __forceinline__ void ola512_walk(int total, const float* __restrict__ taps,
                                            int num_taps, const float2* __restrict__ table,
                                            float2* buf, SegOf seg_of)
{
    const int per = (total + 7) >> 3;
    const int lo = (blockIdx.x & 7) * per;
    const int hi = lo + per < total ? lo + per : total;
    const int step = gridDim.x >> 3;
    int item = lo + static_cast<int>(blockIdx.x >> 3);
    if (item >= hi) {
        return;  // the whole wave
    }
    const int halo = num_taps - 1;
    int t = threadIdx.x;
    Ola512Seg g = seg_of(item);
    float2 next[8];
    ola512_load(g, halo, t, next);
    Ola512Lane s;
    ola512_lane_setup(taps, num_taps, table, t, buf, s);
    while (item < hi) {
        // the lane index through an empty asm: keeps per-lane addresses out of held registers
        asm volatile("" : "+v"(t));
        float2 u[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            u[r] = next[r];
        }
        const Ola512Seg cur = g;
        item += step;
        if (item < hi) {
            g = seg_of(item);
            ola512_load(g, halo, t, next);
        }
        ola512_filter(cur, halo, u, buf, s, t);
    }
}

__global__ __launch_bounds__(k512Lanes) void fir_ola512_kernel(
    const float2* __restrict__ x, const float* __restrict__ taps, float2* __restrict__ y, int n,
    int num_taps, int segments, const float2* __restrict__ table)
{
    __shared__ float2 buf[k512Slots];
    const long long q = kOla512P - (num_taps - 1);
    ola512_walk(segments, taps, num_taps, table, buf,
                [&](int seg) { return Ola512Seg{x, y, n, seg * q}; });
}

// rf_fir_batched: channel blockIdx.y, persistent over its segments.
__global__ __launch_bounds__(k512Lanes) void fir_ola512_batched_kernel(
    FirBatch batch, const float* __restrict__ taps, int num_taps,
    const float2* __restrict__ table)
{
    __shared__ float2 buf[k512Slots];
    const int c = blockIdx.y;
    const int n = batch.n[c];
    const int q = kOla512P - (num_taps - 1);
    const float2* x = batch.x[c];
    float2* y = batch.y[c];
    const int segments = static_cast<int>((static_cast<long long>(n) + q - 1) / q);
    ola512_walk(segments, taps, num_taps, table, buf, [&](int seg) {
        return Ola512Seg{x, y, n, static_cast<long long>(seg) * q};
    });
}

// rf_fir_batch: the {capture, segment} items of all captures.
__global__ __launch_bounds__(k512Lanes) void fir_ola512_batch_kernel(
    const FirCapture* __restrict__ caps, const int2* __restrict__ items, int total,
    const float* __restrict__ taps, int num_taps, const float2* __restrict__ table)
{
    __shared__ float2 buf[k512Slots];
    const long long q = kOla512P - (num_taps - 1);
    ola512_walk(total, taps, num_taps, table, buf, [&](int item) {
        const int2 it = items[item];
        const FirCapture c = caps[it.x];
        return Ola512Seg{c.x, c.y, c.n, it.y * q};
    });
}

// --- host side ------------------------------------------------------------------------------

struct ImplInfo {
    const char* name;
    const char* summary;
};

const ImplInfo kFirImpls[RF_FIR_IMPL_COUNT] = {
    {"naive", "one lane per output sample, global memory only"},
    {"tiled", "samples staged in an LDS tile with a halo"},
    {"tiled_const", "tiled, with the taps read through the scalar (constant) cache"},
    {"register", "4, 8 or 16 outputs per lane from a register window over a padded LDS tile"},
    {"fft_ola", "overlap-save with a 512-point transform per wavefront for 17..225 taps and a "
                "4096-point one for 226..2049 taps; the register kernel outside those ranges"},
    {"mfma_f16x3", "Toeplitz GEMM on FP16 matrix cores, hi/lo split, FP32 accumulate, for "
                   "66..2049 taps; the register kernel outside that range"},
    {"mfma_bf16x3", "Toeplitz GEMM on BF16 matrix cores, hi/lo split, FP32 accumulate, for "
                    "66..2049 taps; the register kernel outside that range"},
};

void check_fir_args(int impl, const void* x, const float* taps, const void* y, int n,
                    int num_taps, int block)
{
    if (impl < 0 || impl >= RF_FIR_IMPL_COUNT) {
        fail(RF_INVALID_ARG, "fir: unknown implementation " + std::to_string(impl));
    }
    if (n < 0) {
        fail(RF_INVALID_ARG, "fir: negative sample count");
    }
    if (num_taps < 1) {
        fail(RF_INVALID_ARG, "fir: need at least one tap");
    }
    if (block < 32 || block > 1024 || block % 32 != 0) {
        fail(RF_INVALID_ARG, "fir: --block needs a multiple of 32 up to 1024, got "
                                 + std::to_string(block));
    }
    if (n > 0 && (x == nullptr || y == nullptr || taps == nullptr)) {
        fail(RF_INVALID_ARG, "fir: NULL buffer");
    }
}

void check_const_taps(int num_taps, const char* impl)
{
    if (num_taps > RF_MAX_CONST_TAPS) {
        fail(RF_LIMIT, std::string("fir ") + impl + ": " + std::to_string(num_taps)
                           + " taps, more than the " + std::to_string(RF_MAX_CONST_TAPS)
                           + " the constant-cache variants accept; use --impl naive");
    }
}

std::size_t tiled_bytes(int num_taps, int block, const char* impl)
{
    const std::size_t bytes = (static_cast<std::size_t>(num_taps) - 1 + block) * sizeof(float2);
    const int limit = max_lds_per_block();
    if (bytes > static_cast<std::size_t>(limit)) {
        fail(RF_LIMIT, std::string("fir ") + impl + ": " + std::to_string(num_taps)
                           + " taps need " + std::to_string(bytes)
                           + " bytes of LDS per workgroup, more than the "
                           + std::to_string(limit) + " available; use --impl naive");
    }
    return bytes;
}

std::size_t register_bytes_for(int r, int num_taps, int block)
{
    const long long tpad = ceil_div(num_taps, r) * r;
    const long long len = tpad + static_cast<long long>(block) * r;
    return static_cast<std::size_t>(len / r * (r + 2)) * sizeof(float2);
}

// Outputs per lane: 16 when the launch still has four or more blocks per CU, else 8 or 4;
// and the largest of those whose tile fits in LDS. (2^20 samples x 63 taps, 256 lanes:
// 4 outputs per lane, 1024 blocks, 10.4 us per call against 11.5 us with 8 and 512 blocks.)
int register_outputs(int num_taps, int block, long long longest, std::size_t* bytes)
{
    const int limit = max_lds_per_block();
    const long long want_blocks = 4LL * cu_count();
    const int candidates[3] = {16, 8, 4};
    int chosen = 0;
    for (int r : candidates) {
        if (register_bytes_for(r, num_taps, block) > static_cast<std::size_t>(limit)) {
            continue;
        }
        chosen = r;
        if (ceil_div(longest, static_cast<long long>(block) * r) >= want_blocks) {
            break;
        }
    }
    if (chosen == 0) {
        fail(RF_LIMIT, "fir register: " + std::to_string(num_taps) + " taps need "
                           + std::to_string(register_bytes_for(4, num_taps, block))
                           + " bytes of LDS per workgroup, more than the "
                           + std::to_string(limit) + " available; use --impl naive");
    }
    *bytes = register_bytes_for(chosen, num_taps, block);
    return chosen;
}

template <class Kernel>
void allow_lds(Kernel kernel, std::size_t bytes)
{
    allow_large_lds(reinterpret_cast<const void*>(kernel), bytes);
}

// `forced_r` (4, 8 or 16) overrides the outputs-per-lane choice; 0 keeps it.
void launch_register(const float2* x, const float* taps, float2* y, int n, int num_taps,
                     int block, hipStream_t stream, int forced_r = 0)
{
    std::size_t bytes = 0;
    int r = register_outputs(num_taps, block, n, &bytes);
    if (forced_r != 0) {
        r = forced_r;
        bytes = register_bytes_for(r, num_taps, block);
        if (bytes > static_cast<std::size_t>(max_lds_per_block())) {
            fail(RF_LIMIT, "fir register: " + std::to_string(r) + " outputs per lane need "
                               + std::to_string(bytes) + " bytes of LDS");
        }
    }
    const unsigned grid = static_cast<unsigned>(ceil_div(n, static_cast<long long>(block) * r));
    if (r == 16) {
        allow_lds(fir_register_kernel<16>, bytes);
        fir_register_kernel<16><<<grid, block, bytes, stream>>>(x, taps, y, n, num_taps);
    } else if (r == 8) {
        allow_lds(fir_register_kernel<8>, bytes);
        fir_register_kernel<8><<<grid, block, bytes, stream>>>(x, taps, y, n, num_taps);
    } else {
        allow_lds(fir_register_kernel<4>, bytes);
        fir_register_kernel<4><<<grid, block, bytes, stream>>>(x, taps, y, n, num_taps);
    }
    check_launch("fir_register_kernel");
}

// The 4096-point overlap-save: from where the 512-point form stops (kOla512MaxTaps, where the
// two measured equal on 2^24 samples) up to kOlaMaxTaps.
bool ola_applies(int num_taps)
{
    return num_taps > kOla512MaxTaps && num_taps <= kOlaMaxTaps;
}

// The matrix-core variants cover the same tap range as overlap-save; shorter filters run the
// direct register kernel (exact for a single tap), longer ones too (their LDS would not fit).
bool mfma_applies(int num_taps)
{
    return num_taps > kOlaMinTaps && num_taps <= kOlaMaxTaps;
}

// Short filters: the 512-point overlap-save (a single tap stays direct, and exact; below
// kOla512MinTaps the direct form is cheaper).
bool ola512_applies(int num_taps)
{
    return num_taps >= kOla512MinTaps && num_taps <= kOla512MaxTaps;
}

// Points of the overlap-save transform fft_ola uses for num_taps taps: 4096, 512, or 0 (the
// register kernel).
int ola_points(int num_taps)
{
    return ola_applies(num_taps) ? kOlaP : (ola512_applies(num_taps) ? kOla512P : 0);
}

// Persistent one-wave workgroups: kOla512WavesPerCu per CU, never more than the items,
// rounded up to a multiple of 8 for the XCD-major walk.
unsigned ola512_grid(long long items)
{
    const long long cap = static_cast<long long>(kOla512WavesPerCu) * cu_count();
    return static_cast<unsigned>(ceil_div(std::min<long long>(std::max<long long>(items, 1), cap),
                                          8LL) * 8);
}

// The whole short-filter overlap-save in one launch (the waves make the tap spectrum).
void launch_ola512(const float2* x, const float* taps, float2* y, int n, int num_taps,
                   hipStream_t stream)
{
    const long long q = kOla512P - (num_taps - 1);
    const long long segments = ceil_div(n, q);
    fir_ola512_kernel<<<ola512_grid(segments), k512Lanes, 0, stream>>>(
        x, taps, y, n, num_taps, static_cast<int>(segments), twiddle_table_4096());
    check_launch("fir_ola512_kernel");
}

void launch_ola_taps(const float* taps, int num_taps, float2* spec, hipStream_t stream)
{
    fir_ola_taps_kernel<<<1, k4096Lanes, 0, stream>>>(taps, num_taps, spec,
                                                      fft4096_stage_table());
    check_launch("fir_ola_taps_kernel");
}

// The overlap-save filtering kernel over x with a ready tap spectrum.
void launch_ola(const float2* x, const float2* spec, float2* y, int n, int num_taps,
                hipStream_t stream)
{
    const long long q = kOlaP - (num_taps - 1);
    const long long segments = ceil_div(n, q);
    const long long per_xcd = ceil_div(segments, 8);
    fir_ola_kernel<<<static_cast<unsigned>(8 * per_xcd), k4096Lanes, 0, stream>>>(
        x, spec, y, n, num_taps, segments, per_xcd, fft4096_stage_table());
    check_launch("fir_ola_kernel");
}


// Largest tap count of each overlap-save length when a configuration forces it (at least 64
// outputs per 512-point segment; the 4096-point form as rf_fir).
constexpr int kOla512LimitTaps = 449;

bool ola_points_fit(int points, int num_taps)
{
    return num_taps >= 1 && num_taps <= (points == kOla512P ? kOla512LimitTaps : kOlaMaxTaps);
}

// fft_ola with a given transform length (tuner configurations): the 512-point form in one
// launch, or the 4096-point form with its tap spectrum in the stream's workspace block.
void launch_ola_points(int points, const float2* x, const float* taps, float2* y, int n,
                       int num_taps, hipStream_t stream)
{
    if (points == kOla512P) {
        launch_ola512(x, taps, y, n, num_taps, stream);
        return;
    }
    float2* h = static_cast<float2*>(stream_workspace(stream));
    launch_ola_taps(taps, num_taps, h, stream);
    launch_ola(x, h, y, n, num_taps, stream);
}

// `workspace`: at least kStreamWorkspaceBytes of device memory for the tap spectrum / split
// taps of fft_ola and the matrix-core variants, or nullptr for the stream's own block.
void fir_launch(int impl, const float2* x, const float* taps, float2* y, int n, int num_taps,
                int block, hipStream_t stream, void* workspace = nullptr)
{
    if (n == 0) {
        return;
    }
    const unsigned grid = static_cast<unsigned>(ceil_div(n, block));
    switch (impl) {
    case RF_FIR_NAIVE:
        fir_naive_kernel<<<grid, block, 0, stream>>>(x, taps, y, n, num_taps);
        check_launch("fir_naive_kernel");
        return;
    case RF_FIR_TILED: {
        const std::size_t bytes = tiled_bytes(num_taps, block, "tiled");
        allow_lds(fir_tiled_kernel<false>, bytes);
        fir_tiled_kernel<false><<<grid, block, bytes, stream>>>(x, taps, y, n, num_taps);
        check_launch("fir_tiled_kernel");
        return;
    }
    case RF_FIR_TILED_CONST: {
        check_const_taps(num_taps, "tiled_const");
        const std::size_t bytes = tiled_bytes(num_taps, block, "tiled_const");
        allow_lds(fir_tiled_kernel<true>, bytes);
        fir_tiled_kernel<true><<<grid, block, bytes, stream>>>(x, taps, y, n, num_taps);
        check_launch("fir_tiled_const_kernel");
        return;
    }
    case RF_FIR_REGISTER:
        check_const_taps(num_taps, "register");
        launch_register(x, taps, y, n, num_taps, block, stream);
        return;
    case RF_FIR_FFT_OLA: {
        if (ola512_applies(num_taps)) {
            launch_ola512(x, taps, y, n, num_taps, stream);
            return;
        }
        if (!ola_applies(num_taps)) {
            launch_register(x, taps, y, n, num_taps, block, stream);
            return;
        }
        float2* h = static_cast<float2*>(workspace != nullptr ? workspace
                                                              : stream_workspace(stream));
        launch_ola_taps(taps, num_taps, h, stream);
        launch_ola(x, h, y, n, num_taps, stream);
        return;
    }
    case RF_FIR_MFMA_F16X3:
    case RF_FIR_MFMA_BF16X3:
        if (!mfma_applies(num_taps)) {
            launch_register(x, taps, y, n, num_taps, block, stream);
            return;
        }
        fir_mfma_launch(impl == RF_FIR_MFMA_BF16X3, x, taps, y, n, num_taps, stream,
                        workspace != nullptr ? workspace : stream_workspace(stream));
        return;
    }
}

void fir_batched_launch(int impl, const FirBatch& batch, int count, int longest,
                        const float* taps, int num_taps, hipStream_t stream)
{
    const int block = RF_DEFAULT_FIR_BLOCK;
    if (impl == RF_FIR_MFMA_F16X3 || impl == RF_FIR_MFMA_BF16X3) {
        // One launch per channel: the matrix-core kernel's tile shape is per capture.
        void* ws = stream_workspace(stream);
        for (int c = 0; c < count; ++c) {
            fir_launch(impl, batch.x[c], taps, batch.y[c], batch.n[c], num_taps, block, stream,
                       ws);
        }
        return;
    }
    switch (impl) {
    case RF_FIR_NAIVE: {
        const dim3 grid(static_cast<unsigned>(ceil_div(longest, block)), count);
        fir_naive_batched_kernel<<<grid, block, 0, stream>>>(batch, taps, num_taps);
        check_launch("fir_naive_batched_kernel");
        return;
    }
    case RF_FIR_TILED:
    case RF_FIR_TILED_CONST: {
        const bool ct = impl == RF_FIR_TILED_CONST;
        if (ct) {
            check_const_taps(num_taps, "tiled_const");
        }
        const std::size_t bytes = tiled_bytes(num_taps, block, ct ? "tiled_const" : "tiled");
        const dim3 grid(static_cast<unsigned>(ceil_div(longest, block)), count);
        if (ct) {
            allow_lds(fir_tiled_batched_kernel<true>, bytes);
            fir_tiled_batched_kernel<true><<<grid, block, bytes, stream>>>(batch, taps, num_taps);
        } else {
            allow_lds(fir_tiled_batched_kernel<false>, bytes);
            fir_tiled_batched_kernel<false><<<grid, block, bytes, stream>>>(batch, taps,
                                                                           num_taps);
        }
        check_launch("fir_tiled_batched_kernel");
        return;
    }
    case RF_FIR_REGISTER:
    case RF_FIR_FFT_OLA: {
        if (impl == RF_FIR_FFT_OLA && ola_applies(num_taps)) {
            float2* spec = static_cast<float2*>(stream_workspace(stream));
            launch_ola_taps(taps, num_taps, spec, stream);
            const long long q = kOlaP - (num_taps - 1);
            const dim3 grid(static_cast<unsigned>(ceil_div(longest, q)), count);
            fir_ola_batched_kernel<<<grid, k4096Lanes, 0, stream>>>(
                batch, spec, num_taps, fft4096_stage_table());
            check_launch("fir_ola_batched_kernel");
            return;
        }
        if (impl == RF_FIR_FFT_OLA && ola512_applies(num_taps)) {
            const long long q = kOla512P - (num_taps - 1);
            const dim3 grid(ola512_grid(ceil_div(longest, q)), count);
            fir_ola512_batched_kernel<<<grid, k512Lanes, 0, stream>>>(batch, taps, num_taps,
                                                                       twiddle_table_4096());
            check_launch("fir_ola512_batched_kernel");
            return;
        }
        if (impl == RF_FIR_REGISTER) {
            check_const_taps(num_taps, "register");
        }
        std::size_t bytes = 0;
        const int r = register_outputs(num_taps, block, longest, &bytes);
        const dim3 grid(static_cast<unsigned>(ceil_div(longest, static_cast<long long>(block) * r)),
                        count);
        if (r == 16) {
            allow_lds(fir_register_batched_kernel<16>, bytes);
            fir_register_batched_kernel<16><<<grid, block, bytes, stream>>>(batch, taps, num_taps);
        } else if (r == 8) {
            allow_lds(fir_register_batched_kernel<8>, bytes);
            fir_register_batched_kernel<8><<<grid, block, bytes, stream>>>(batch, taps, num_taps);
        } else {
            allow_lds(fir_register_batched_kernel<4>, bytes);
            fir_register_batched_kernel<4><<<grid, block, bytes, stream>>>(batch, taps, num_taps);
        }
        check_launch("fir_register_batched_kernel");
        return;
    }
    }
}

// Device memory a one-shot call of `impl` with num_taps taps needs beside its inputs and
// outputs: the tap spectrum (fft_ola) or the split taps (MFMA variants); 0 for the direct
// kernels.
std::size_t fir_workspace_bytes(int impl, int num_taps)
{
    if (impl == RF_FIR_FFT_OLA && ola_applies(num_taps)) {
        return kOlaP * sizeof(float2);
    }
    if ((impl == RF_FIR_MFMA_F16X3 || impl == RF_FIR_MFMA_BF16X3) && mfma_applies(num_taps)) {
        return fir_mfma_workspace_bytes(num_taps);
    }
    return 0;
}

}  // namespace
}  // namespace rf

// Scratch of rf_fir_noalloc: the device memory rf_fir takes from a stream's workspace block,
// here a fixed array of the library's code object, so that no call allocates. Each GPU holds
// its own copy for the life of the process; every rf_fir_noalloc call on a GPU shares it.
constexpr std::size_t kNoAllocScratchBytes = rf::kStreamWorkspaceBytes;  // 72 KiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace rf {
namespace {

constexpr int kScratchDevices = 64;
std::once_flag g_scratch_once[kScratchDevices];
void* g_scratch[kScratchDevices] = {};

// Device address of `scratch` on the current device, looked up once per device (the name
// itself is only a host-side stand-in). The lookup allocates nothing.
void* noalloc_scratch()
{
    int device = 0;
    RF_HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kScratchDevices) {
        fail(RF_LIMIT, "fir noalloc: device index out of range");
    }
    std::call_once(g_scratch_once[device], [device] {
        void* p = nullptr;
        RF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = p;
    });
    if (g_scratch[device] == nullptr) {
        fail(RF_HIP_ERROR, "fir noalloc: scratch lookup failed on an earlier call");
    }
    return g_scratch[device];
}

}  // namespace
}  // namespace rf

using rf::guarded;

extern "C" {

int rf_fir_impl_count(void)
{
    return RF_FIR_IMPL_COUNT;
}

const char* rf_fir_impl_name(int impl)
{
    return (impl >= 0 && impl < RF_FIR_IMPL_COUNT) ? rf::kFirImpls[impl].name : nullptr;
}

const char* rf_fir_impl_summary(int impl)
{
    return (impl >= 0 && impl < RF_FIR_IMPL_COUNT) ? rf::kFirImpls[impl].summary : nullptr;
}

rf_status rf_fir_parse_impl(const char* name, int* impl)
{
    return guarded([&] {
        if (name == nullptr || impl == nullptr) {
            rf::fail(RF_INVALID_ARG, "fir: NULL implementation name");
        }
        for (int i = 0; i < RF_FIR_IMPL_COUNT; ++i) {
            if (std::strcmp(name, rf::kFirImpls[i].name) == 0) {
                *impl = i;
                return;
            }
        }
        rf::fail(RF_INVALID_ARG, std::string("unknown fir implementation '") + name + "'");
    });
}

const char* rf_fir_kernel_name(int impl, int num_taps)
{
    if (impl < 0 || impl >= RF_FIR_IMPL_COUNT || num_taps < 1) {
        return nullptr;
    }
    switch (impl) {
    case RF_FIR_NAIVE:
        return "fir_naive_kernel";
    case RF_FIR_TILED:
    case RF_FIR_TILED_CONST:
        return "fir_tiled_kernel";
    case RF_FIR_FFT_OLA:
        return rf::ola_applies(num_taps)
                   ? "fir_ola_kernel"
                   : (rf::ola512_applies(num_taps) ? "fir_ola512_kernel" : "fir_register_kernel");
    case RF_FIR_MFMA_F16X3:
    case RF_FIR_MFMA_BF16X3:
        return rf::mfma_applies(num_taps) ? "fir_mfma_kernel" : "fir_register_kernel";
    default:
        return "fir_register_kernel";
    }
}

rf_status rf_fir(int impl, const rf_cf32* x, const float* taps, rf_cf32* y, int num_samples,
                 int num_taps, int block, rf_stream stream)
{
    return guarded([&] {
        const int b = block == 0 ? RF_DEFAULT_FIR_BLOCK : block;
        rf::check_fir_args(impl, x, taps, y, num_samples, num_taps, b);
        rf::fir_launch(impl, reinterpret_cast<const float2*>(x), taps,
                       reinterpret_cast<float2*>(y), num_samples, num_taps, b,
                       rf::as_stream(stream));
    });
}

rf_status rf_fir_noalloc(int impl, const rf_cf32* x, const float* taps, rf_cf32* y,
                         int num_samples, int num_taps, int block, rf_stream stream)
{
    return guarded([&] {
        const int b = block == 0 ? RF_DEFAULT_FIR_BLOCK : block;
        rf::check_fir_args(impl, x, taps, y, num_samples, num_taps, b);
        const std::size_t need = rf::fir_workspace_bytes(impl, num_taps);
        if (need > kNoAllocScratchBytes) {
            rf::fail(RF_LIMIT, "fir noalloc: " + std::to_string(num_taps) + " taps need "
                                   + std::to_string(need) + " bytes of scratch, more than the "
                                   + std::to_string(kNoAllocScratchBytes) + " it has");
        }
        if (num_samples == 0) {
            return;
        }
        rf::fir_launch(impl, reinterpret_cast<const float2*>(x), taps,
                       reinterpret_cast<float2*>(y), num_samples, num_taps, b,
                       rf::as_stream(stream), need > 0 ? rf::noalloc_scratch() : nullptr);
    });
}

rf_status rf_fir_run_config(const rf_fir_config* config, const rf_cf32* x, const float* taps,
                            rf_cf32* y, int num_samples, int num_taps, rf_stream stream)
{
    return guarded([&] {
        if (config == nullptr) rf::fail(RF_INVALID_ARG, "fir config: NULL");
        const int r = config->outputs_per_lane;
        if (r != 0 && r != 4 && r != 8 && r != 16) {
            rf::fail(RF_INVALID_ARG, "fir config: outputs_per_lane must be 0, 4, 8 or 16");
        }
        rf::check_fir_args(config->impl, x, taps, y, num_samples, num_taps, config->block);
        const int points = config->transform_points;
        if (points != 0 && points != rf::kOla512P && points != rf::kOlaP) {
            rf::fail(RF_INVALID_ARG, "fir config: transform_points must be 0, 512 or 4096");
        }
        if (points != 0 && config->impl != RF_FIR_FFT_OLA) {
            rf::fail(RF_INVALID_ARG, "fir config: transform_points applies to fft_ola only");
        }
        if (points != 0 && !rf::ola_points_fit(points, num_taps)) {
            rf::fail(RF_LIMIT, "fir config: a " + std::to_string(points)
                                   + "-point overlap-save takes 1.."
                                   + std::to_string(points == rf::kOla512P ? rf::kOla512LimitTaps
                                                                           : rf::kOlaMaxTaps)
                                   + " taps, not " + std::to_string(num_taps));
        }
        if (num_samples == 0) {
            return;
        }
        if (points != 0) {
            rf::launch_ola_points(points, reinterpret_cast<const float2*>(x), taps,
                                  reinterpret_cast<float2*>(y), num_samples, num_taps,
                                  rf::as_stream(stream));
            return;
        }
        if (config->impl == RF_FIR_REGISTER && r != 0) {
            rf::launch_register(reinterpret_cast<const float2*>(x), taps,
                                reinterpret_cast<float2*>(y), num_samples, num_taps,
                                config->block, rf::as_stream(stream), r);
            return;
        }
        rf::fir_launch(config->impl, reinterpret_cast<const float2*>(x), taps,
                       reinterpret_cast<float2*>(y), num_samples, num_taps, config->block,
                       rf::as_stream(stream));
    });
}

rf_status rf_fir_tune(const rf_cf32* x, const float* taps, rf_cf32* y, int num_samples,
                      int num_taps, unsigned impl_mask, rf_stream stream, rf_fir_config* best)
{
    return guarded([&] {
        if (best == nullptr) rf::fail(RF_INVALID_ARG, "fir tune: NULL result");
        rf::check_fir_args(RF_FIR_REGISTER, x, taps, y, num_samples, num_taps,
                           RF_DEFAULT_FIR_BLOCK);
        if ((impl_mask & ((1u << RF_FIR_IMPL_COUNT) - 1)) == 0) {
            rf::fail(RF_INVALID_ARG, "fir tune: no variant selected");
        }
        std::vector<rf_fir_config> candidates;
        for (int impl = 0; impl < RF_FIR_IMPL_COUNT; ++impl) {
            if ((impl_mask >> impl & 1u) == 0) {
                continue;
            }
            if (impl == RF_FIR_TILED_CONST && num_taps > RF_MAX_CONST_TAPS) {
                continue;
            }
            if (impl == RF_FIR_REGISTER) {
                for (int block : {64, 128, 256, 512}) {
                    for (int r : {4, 8, 16}) {
                        if (rf::register_bytes_for(r, num_taps, block)
                            <= static_cast<std::size_t>(rf::max_lds_per_block())) {
                            candidates.push_back(rf_fir_config{impl, block, r, 0.0f, 0});
                        }
                    }
                }
            } else if (impl == RF_FIR_NAIVE || impl == RF_FIR_TILED || impl == RF_FIR_TILED_CONST) {
                for (int block : {128, 256, 512}) {
                    candidates.push_back(rf_fir_config{impl, block, 0, 0.0f, 0});
                }
            } else if (impl == RF_FIR_FFT_OLA) {
                // the transform length: each form the tap count allows (a single tap, or more
                // than 2049, leaves only the library's choice, the register kernel)
                bool any = false;
                if (num_taps >= 2) {
                    for (int points : {rf::kOla512P, rf::kOlaP}) {
                        if (rf::ola_points_fit(points, num_taps)) {
                            candidates.push_back(
                                rf_fir_config{impl, RF_DEFAULT_FIR_BLOCK, 0, 0.0f, points});
                            any = true;
                        }
                    }
                }
                if (!any) {
                    candidates.push_back(rf_fir_config{impl, RF_DEFAULT_FIR_BLOCK, 0, 0.0f, 0});
                }
            } else {
                candidates.push_back(rf_fir_config{impl, RF_DEFAULT_FIR_BLOCK, 0, 0.0f, 0});
            }
        }
        hipStream_t s = rf::as_stream(stream);
        hipEvent_t start = nullptr;
        hipEvent_t stop = nullptr;
        RF_HIP_CHECK(hipEventCreate(&start));
        RF_HIP_CHECK(hipEventCreate(&stop));
        bool found = false;
        try {
            for (rf_fir_config& c : candidates) {
                // A candidate that does not fit (LDS) is skipped, not an error of the tuner.
                if (rf_fir_run_config(&c, x, taps, y, num_samples, num_taps, stream) != RF_OK) {
                    continue;
                }
                RF_HIP_CHECK(hipEventRecord(start, s));
                for (int rep = 0; rep < 3; ++rep) {
                    const rf_status st = rf_fir_run_config(&c, x, taps, y, num_samples, num_taps,
                                                           stream);
                    if (st != RF_OK) {
                        rf::fail(st, std::string("fir tune: ") + rf_last_error());
                    }
                }
                RF_HIP_CHECK(hipEventRecord(stop, s));
                RF_HIP_CHECK(hipEventSynchronize(stop));
                float ms = 0.0f;
                RF_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
                c.ms = ms / 3.0f;
                if (!found || c.ms < best->ms) {
                    *best = c;
                    found = true;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(start);
            (void)hipEventDestroy(stop);
            throw;
        }
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        if (!found) rf::fail(RF_LIMIT, "fir tune: no selected variant fits this filter");
    });
}

rf_status rf_fir_batched(int impl, const rf_cf32* const* x, rf_cf32* const* y,
                         const int* num_samples, int count, const float* taps, int num_taps,
                         rf_stream stream)
{
    return guarded([&] {
        if (count < 1 || count > RF_MAX_FIR_BATCH) {
            rf::fail(RF_LIMIT, "batched fir: " + std::to_string(count)
                                   + " channels, expected 1 to "
                                   + std::to_string(RF_MAX_FIR_BATCH));
        }
        if (x == nullptr || y == nullptr || num_samples == nullptr) {
            rf::fail(RF_INVALID_ARG, "batched fir: NULL channel arrays");
        }
        rf::FirBatch batch{};
        int longest = 0;
        for (int c = 0; c < count; ++c) {
            rf::check_fir_args(impl, x[c], taps, y[c], num_samples[c], num_taps,
                               RF_DEFAULT_FIR_BLOCK);
            batch.x[c] = reinterpret_cast<const float2*>(x[c]);
            batch.y[c] = reinterpret_cast<float2*>(y[c]);
            batch.n[c] = num_samples[c];
            longest = num_samples[c] > longest ? num_samples[c] : longest;
        }
        if (longest == 0) {
            return;
        }
        rf::fir_batched_launch(impl, batch, count, longest, taps, num_taps,
                               rf::as_stream(stream));
    });
}

}  // extern "C"

// A filter prepared once for many calls: device copies of the taps and, when the 4096-point
// overlap-save applies, their spectrum (made by the same kernel rf_fir runs per call, so the
// results are the same bits; the 512-point form makes its spectrum inside its one launch).
struct rf_fir_plan {
    int impl = 0;
    int num_taps = 0;
    int block = RF_DEFAULT_FIR_BLOCK;
    float* taps = nullptr;
    float2* spec = nullptr;
};

namespace {

void destroy_fir_plan(rf_fir_plan* plan)
{
    if (plan == nullptr) {
        return;
    }
    if (plan->taps != nullptr) {
        (void)hipFree(plan->taps);
    }
    if (plan->spec != nullptr) {
        (void)hipFree(plan->spec);
    }
    delete plan;
}

}  // namespace

extern "C" {

rf_status rf_fir_plan_create(int impl, const float* taps, int num_taps, int block,
                             rf_stream stream, rf_fir_plan** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "fir plan: NULL output handle");
        *out = nullptr;
        const int b = block == 0 ? RF_DEFAULT_FIR_BLOCK : block;
        // The sample buffers come with each run; check the filter half now.
        static const float2 probe{};
        rf::check_fir_args(impl, &probe, taps, &probe, 1, num_taps, b);
        if (impl == RF_FIR_TILED_CONST || impl == RF_FIR_REGISTER) {
            rf::check_const_taps(num_taps, impl == RF_FIR_REGISTER ? "register" : "tiled_const");
        }
        if (impl == RF_FIR_TILED || impl == RF_FIR_TILED_CONST) {
            (void)rf::tiled_bytes(num_taps, b, impl == RF_FIR_TILED ? "tiled" : "tiled_const");
        }
        hipStream_t s = rf::as_stream(stream);
        rf_fir_plan* plan = new rf_fir_plan;
        plan->impl = impl;
        plan->num_taps = num_taps;
        plan->block = b;
        try {
            RF_HIP_CHECK(hipMalloc(&plan->taps, sizeof(float) * num_taps));
            RF_HIP_CHECK(hipMemcpyAsync(plan->taps, taps, sizeof(float) * num_taps,
                                        hipMemcpyDeviceToDevice, s));
            if (impl == RF_FIR_FFT_OLA && rf::ola_applies(num_taps)) {
                RF_HIP_CHECK(hipMalloc(&plan->spec, sizeof(float2) * rf::kOlaP));
                rf::launch_ola_taps(plan->taps, num_taps, plan->spec, s);
            }
            // Ready for any stream once create returns.
            RF_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            destroy_fir_plan(plan);
            throw;
        }
        *out = plan;
    });
}

rf_status rf_fir_plan_run(const rf_fir_plan* plan, const rf_cf32* x, rf_cf32* y,
                          int num_samples, rf_stream stream)
{
    return guarded([&] {
        if (plan == nullptr) rf::fail(RF_INVALID_ARG, "fir plan: NULL plan");
        rf::check_fir_args(plan->impl, x, plan->taps, y, num_samples, plan->num_taps,
                           plan->block);
        if (num_samples == 0) {
            return;
        }
        const float2* xs = reinterpret_cast<const float2*>(x);
        float2* ys = reinterpret_cast<float2*>(y);
        if (plan->spec != nullptr) {
            rf::launch_ola(xs, plan->spec, ys, num_samples, plan->num_taps, rf::as_stream(stream));
        } else {
            rf::fir_launch(plan->impl, xs, plan->taps, ys, num_samples, plan->num_taps,
                           plan->block, rf::as_stream(stream));
        }
    });
}

rf_status rf_fir_plan_destroy(rf_fir_plan* plan)
{
    return guarded([&] { destroy_fir_plan(plan); });
}

}  // extern "C"

// A persistent FIR over many captures of uneven length sharing one filter (rf_fir_batch_*).
struct rf_fir_batch {
    int impl = 0;
    int num_taps = 0;
    int count = 0;
    int total = 0;                 // work items
    int points = 0;                // overlap-save segments of this transform length, else
                                   // register blocks (0)
    int r = 4;                     // register: outputs per lane
    std::size_t lds = 0;           // register: dynamic LDS per workgroup
    float* taps = nullptr;         // device copy
    float2* spec = nullptr;        // overlap-save tap spectrum
    rf::FirCapture* caps = nullptr;
    int2* items = nullptr;
    std::vector<int> n;
    bool bound = false;
};

namespace {

void destroy_fir_batch(rf_fir_batch* b)
{
    if (b == nullptr) {
        return;
    }
    (void)hipFree(b->taps);
    (void)hipFree(b->spec);
    (void)hipFree(b->caps);
    (void)hipFree(b->items);
    delete b;
}

}  // namespace

extern "C" {

rf_status rf_fir_batch_create(int impl, const float* taps, int num_taps, int count,
                              const int* num_samples, rf_stream stream, rf_fir_batch** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "fir batch: NULL output handle");
        *out = nullptr;
        if (impl != RF_FIR_REGISTER && impl != RF_FIR_FFT_OLA) {
            rf::fail(RF_INVALID_ARG, "fir batch: impl must be register or fft_ola");
        }
        if (count < 1 || num_samples == nullptr) {
            rf::fail(RF_INVALID_ARG, "fir batch: need at least one capture");
        }
        static const float2 probe{};
        rf::check_fir_args(impl, &probe, taps, &probe, 1, num_taps, RF_DEFAULT_FIR_BLOCK);
        const int points = impl == RF_FIR_FFT_OLA ? rf::ola_points(num_taps) : 0;
        if (points == 0) {
            rf::check_const_taps(num_taps, "register");
        }
        long long longest = 0;
        for (int c = 0; c < count; ++c) {
            if (num_samples[c] < 0) rf::fail(RF_INVALID_ARG, "fir batch: negative sample count");
            longest = std::max<long long>(longest, num_samples[c]);
        }
        rf_fir_batch* b = new rf_fir_batch;
        b->impl = impl;
        b->num_taps = num_taps;
        b->count = count;
        b->points = points;
        b->n.assign(num_samples, num_samples + count);
        try {
            long long per;  // outputs per item
            if (points != 0) {
                per = points - (num_taps - 1);
            } else {
                b->r = rf::register_outputs(num_taps, RF_DEFAULT_FIR_BLOCK, longest, &b->lds);
                per = static_cast<long long>(RF_DEFAULT_FIR_BLOCK) * b->r;
            }
            std::vector<int2> items;
            for (int c = 0; c < count; ++c) {
                const long long k = rf::ceil_div(b->n[static_cast<std::size_t>(c)], per);
                for (long long i = 0; i < k; ++i) {
                    items.push_back(make_int2(c, static_cast<int>(i)));
                }
            }
            if (items.size() > static_cast<std::size_t>(INT32_MAX)) {
                rf::fail(RF_LIMIT, "fir batch: more than 2^31 work items");
            }
            b->total = static_cast<int>(items.size());
            hipStream_t s = rf::as_stream(stream);
            RF_HIP_CHECK(hipMalloc(&b->taps, sizeof(float) * num_taps));
            RF_HIP_CHECK(hipMemcpyAsync(b->taps, taps, sizeof(float) * num_taps,
                                        hipMemcpyDeviceToDevice, s));
            if (points == rf::kOlaP) {
                RF_HIP_CHECK(hipMalloc(&b->spec, sizeof(float2) * rf::kOlaP));
                rf::launch_ola_taps(b->taps, num_taps, b->spec, s);
            }
            RF_HIP_CHECK(hipMalloc(&b->caps, sizeof(rf::FirCapture) * count));
            if (!items.empty()) {
                RF_HIP_CHECK(hipMalloc(&b->items, sizeof(int2) * items.size()));
                RF_HIP_CHECK(hipMemcpyAsync(b->items, items.data(), sizeof(int2) * items.size(),
                                            hipMemcpyHostToDevice, s));
            }
            RF_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            destroy_fir_batch(b);
            throw;
        }
        *out = b;
    });
}

rf_status rf_fir_batch_bind(rf_fir_batch* batch, const rf_cf32* const* x, rf_cf32* const* y)
{
    return guarded([&] {
        if (batch == nullptr || x == nullptr || y == nullptr) {
            rf::fail(RF_INVALID_ARG, "fir batch: NULL batch or capture arrays");
        }
        std::vector<rf::FirCapture> caps(static_cast<std::size_t>(batch->count));
        for (int c = 0; c < batch->count; ++c) {
            const int n = batch->n[static_cast<std::size_t>(c)];
            if (n > 0 && (x[c] == nullptr || y[c] == nullptr)) {
                rf::fail(RF_INVALID_ARG, "fir batch: NULL buffer for capture " + std::to_string(c));
            }
            caps[static_cast<std::size_t>(c)] =
                rf::FirCapture{reinterpret_cast<const float2*>(x[c]), reinterpret_cast<float2*>(y[c]), n};
        }
        RF_HIP_CHECK(hipMemcpy(batch->caps, caps.data(), sizeof(rf::FirCapture) * caps.size(),
                               hipMemcpyHostToDevice));
        batch->bound = true;
    });
}

rf_status rf_fir_batch_run(const rf_fir_batch* batch, rf_stream stream)
{
    return guarded([&] {
        if (batch == nullptr) rf::fail(RF_INVALID_ARG, "fir batch: NULL batch");
        if (!batch->bound) rf::fail(RF_INVALID_ARG, "fir batch: run before bind");
        if (batch->total == 0) {
            return;
        }
        hipStream_t s = rf::as_stream(stream);
        // persistent grid: four workgroups per CU (a multiple of 8 for the XCD-major walk)
        const long long cap = 4LL * rf::cu_count();
        const unsigned grid = static_cast<unsigned>(
            rf::ceil_div(std::min<long long>(batch->total, cap), 8) * 8);
        if (batch->points == rf::kOla512P) {
            rf::fir_ola512_batch_kernel<<<rf::ola512_grid(batch->total), rf::k512Lanes, 0, s>>>(
                batch->caps, batch->items, batch->total, batch->taps, batch->num_taps,
                rf::twiddle_table_4096());
            rf::check_launch("fir_ola512_batch_kernel");
            return;
        }
        if (batch->points == rf::kOlaP) {
            rf::fir_ola_batch_kernel<<<grid, rf::k4096Lanes, 0, s>>>(
                batch->caps, batch->items, batch->total, batch->spec, batch->num_taps,
                rf::fft4096_stage_table());
            rf::check_launch("fir_ola_batch_kernel");
            return;
        }
        auto launch = [&](auto kernel) {
            rf::allow_large_lds(reinterpret_cast<const void*>(kernel), batch->lds);
            kernel<<<grid, RF_DEFAULT_FIR_BLOCK, batch->lds, s>>>(
                batch->caps, batch->items, batch->total, batch->taps, batch->num_taps);
        };
        if (batch->r == 16) {
            launch(rf::fir_register_batch_kernel<16>);
        } else if (batch->r == 8) {
            launch(rf::fir_register_batch_kernel<8>);
        } else {
            launch(rf::fir_register_batch_kernel<4>);
        }
        rf::check_launch("fir_register_batch_kernel");
    });
}

rf_status rf_fir_batch_destroy(rf_fir_batch* batch)
{
    return guarded([&] { destroy_fir_batch(batch); });
}

}  // extern "C"
