// 4096-point transforms held in registers between three radix-16 Stockham passes (spans 1, 16
// and 256), 256 lanes with 16 values each. Used by the overlap-save FIR and by the row pass of
// the cross-correlations, whose rows are 4096 points long.
//
// A lane starts with the inputs t + 256 r (r < 16) in registers and ends with the outputs
// t + 256 r in registers, so a caller can read its inputs straight from HBM, apply a pointwise
// operation between two transforms without touching LDS, and write its outputs straight back;
// each transform makes two LDS round trips. LDS slot of element j: j + j/16, which puts the 16
// lanes of each stride-16 write and of each (t/16)*256 + t%16 write on 16 different banks.
// Twiddles: per-device table [span][k][r] = exp(-2 pi i r k / (16 span)), r < 16, k < span,
// span 16 at offset 0 and span 256 at offset 256 (fft4096_stage_table()); a lane reads its 16
// factors as 8 16-byte loads.
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

template <bool INV>
__device__ __forceinline__ void twiddle_dft16(float2* u, int k, const float2* stage)
{
    const float4* tw = reinterpret_cast<const float4*>(stage + 16 * k);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float4 w = tw[q];
        if (q > 0) {
            u[2 * q] = INV ? cmulconj(u[2 * q], make_float2(w.x, w.y))
                           : cmul(u[2 * q], make_float2(w.x, w.y));
        }
        u[2 * q + 1] = INV ? cmulconj(u[2 * q + 1], make_float2(w.z, w.w))
                           : cmul(u[2 * q + 1], make_float2(w.z, w.w));
    }
    dft16<INV>(u);
}

// Span-1 pass from registers (inputs t + 256 r) into LDS at 16 t + r.
template <bool INV>
__device__ __forceinline__ void fft4096_first(float2* buf, int t, float2* u)
{
    dft16<INV>(u);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        buf[pad16(16 * t + r)] = u[r];
    }
    __syncthreads();
}

// Span-16 pass through LDS.
template <bool INV>
__device__ __forceinline__ void fft4096_middle(float2* buf, int t, const float2* stages)
{
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        u[r] = buf[pad16(t + r * k4096Lanes)];
    }
    __syncthreads();
    const int k = t & 15;
    twiddle_dft16<INV>(u, k, stages);
    const int j = (t - k) * 16 + k;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        buf[pad16(j + 16 * r)] = u[r];
    }
    __syncthreads();
}

// Span-256 pass from LDS; outputs t + 256 r stay in registers.
template <bool INV>
__device__ __forceinline__ void fft4096_last(const float2* buf, int t, float2* u,
                                             const float2* stages)
{
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        u[r] = buf[pad16(t + r * k4096Lanes)];
    }
    twiddle_dft16<INV>(u, t, stages + 256);
}

// The whole transform, unnormalized: u[r] = x[t + 256 r] in, X[t + 256 r] out. `buf` holds
// k4096Slots values; the caller separates two uses of the same buffer with a barrier once
// every lane has read what the previous transform left there.
template <bool INV>
__device__ __forceinline__ void fft4096(float2* buf, int t, float2* u, const float2* stages)
{
    fft4096_first<INV>(buf, t, u);
    fft4096_middle<INV>(buf, t, stages);
    fft4096_last<INV>(buf, t, u, stages);
}

}  // namespace rf
