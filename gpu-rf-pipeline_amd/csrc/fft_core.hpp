// Building blocks of the library's own FFT (no vendor FFT library on the hot path).
//
// Method: radix-8/4/2 Stockham autosort passes (Stockham 1966; the per-pass index map is the
// one in E. Bainville, "OpenCL Fast Fourier Transform", 2010). A transform of P = 2^m points
// runs as passes of radix R with a running span p (1, R0, R0*R1, ...). Butterfly i of a pass:
//   k = i mod p
//   u[r] = src[i + r * P/R] * exp(-+2 pi i r k / (p R))     r = 0 .. R-1
//   u    = DFT_R(u)
//   dst[(i - k) * R + k + r * p] = u[r]
// Input and output are both in natural order after the last pass.
//
// Twiddles come from a float32 table of exp(-2 pi i m / 4096), m < 4096, built in float64 on
// the host (tables are the only place float64 is used). A transform of P <= 4096 points reads
// it with stride 4096 / P. Longer transforms (four-step) use two-level tables per plan.
//
// LDS layout: element j is stored at j + (j >> 3) (one pad slot per 8 complex values), which
// spreads the stride-R writes of the first passes over distinct banks.
#pragma once

#include "rf_common.hpp"

namespace rf {

constexpr int kTwiddleLog2 = 12;
constexpr int kTwiddleSize = 1 << kTwiddleLog2;  // 4096
constexpr int kMaxLdsFft = 4096;                 // largest single-workgroup transform

__host__ __device__ constexpr int lds_pad(int j)
{
    return j + (j >> 3);
}

// LDS slots (float2) for a padded transform of P points.
__host__ __device__ constexpr int lds_slots(int P)
{
    return P + (P >> 3) + 1;
}

// Radix of pass `pass` for a transform of 2^log2P points: radix-8 first, then one radix-4 or
// radix-2 pass for the remainder.
__host__ __device__ __forceinline__ int pass_radix_log2(int log2P, int pass)
{
    const int full8 = log2P / 3;
    return pass < full8 ? 3 : (log2P - 3 * full8);
}

__host__ __device__ __forceinline__ int pass_count(int log2P)
{
    return log2P / 3 + ((log2P % 3) ? 1 : 0);
}

// Forward twiddle exp(-2 pi i idx / 4096) from the table, conjugated for the inverse.
template <bool INV>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ table, int idx)
{
    const float2 w = table[idx];
    return INV ? make_float2(w.x, -w.y) : w;
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b)
{
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// 4-point DFT, forward kernel exp(-2 pi i / 4) = -i.
template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3)
{
    const float2 s02 = cadd(v0, v2);
    const float2 d02 = csub(v0, v2);
    const float2 s13 = cadd(v1, v3);
    const float2 d13 = csub(v1, v3);
    // forward: -i * d13 = (d13.y, -d13.x); inverse: +i * d13 = (-d13.y, d13.x)
    const float2 m = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
    v0 = cadd(s02, s13);
    v2 = csub(s02, s13);
    v1 = cadd(d02, m);
    v3 = csub(d02, m);
}

// 8-point DFT as two 4-point DFTs of the even and odd samples and a radix-2 combine.
template <bool INV>
__device__ __forceinline__ void dft8(float2* v)
{
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float c = 0.70710678118654752440f;
    // o_k *= exp(-+2 pi i k / 8)
    float2 t1, t2, t3;
    if (!INV) {
        t1 = make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));   // * (c - ic)
        t2 = make_float2(o2.y, -o2.x);                            // * (-i)
        t3 = make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));  // * (-c - ic)
    } else {
        t1 = make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y));   // * (c + ic)
        t2 = make_float2(-o2.y, o2.x);                            // * (+i)
        t3 = make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y));  // * (-c + ic)
    }
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1);
    v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2);
    v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3);
    v[7] = csub(e3, t3);
}

// One radix-2^lr butterfly on u[0 .. R-1] with its Stockham twiddles. `k` = i mod p and
// `tw_step` = 4096 / (p R): u[r] *= W^(r k tw_step).
template <bool INV>
__device__ __forceinline__ void twiddle_and_dft(float2* u, int lr, int k,
                                                const float2* __restrict__ table, int tw_step)
{
    const int R = 1 << lr;
    if (k != 0) {
        const int base = k * tw_step;
#pragma unroll
        for (int r = 1; r < 8; ++r) {
            if (r < R) {
                u[r] = cmul(u[r], twiddle<INV>(table, r * base));
            }
        }
    }
    if (lr == 3) {
        dft8<INV>(u);
    } else if (lr == 2) {
        dft4<INV>(u[0], u[1], u[2], u[3]);
    } else if (lr == 1) {
        dft2<INV>(u[0], u[1]);
    }
}

// A transform of P = 2^log2P (8 <= P <= 4096) points done by P/8 threads of a workgroup on
// data already in LDS (padded layout), in place, natural order in and out. Each thread reads
// eight values per pass (one radix-8 butterfly, two radix-4 or four radix-2 butterflies),
// so the pass needs one barrier between the reads and the writes. `t` is the thread's index
// within the transform; every thread of the workgroup must call this (it has barriers).
// `active` is false for threads that only take part in the barriers.
template <bool INV>
__device__ void lds_fft(float2* buf, int log2P, int t, bool active,
                        const float2* __restrict__ table)
{
    const int P = 1 << log2P;
    const int eighth = P >> 3;
    const int passes = pass_count(log2P);
    int p = 1;
    for (int pass = 0; pass < passes; ++pass) {
        const int lr = pass_radix_log2(log2P, pass);
        const int R = 1 << lr;
        const int nb = 8 >> lr;       // butterflies per thread
        const int stride = P >> lr;   // P / R
        float2 u[8];
        if (active) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int bi = b >> lr;          // which butterfly
                const int r = b & (R - 1);       // which input of it
                const int i = t + bi * eighth;   // butterfly index
                if (bi < nb) {
                    u[b] = buf[lds_pad(i + r * stride)];
                }
            }
        }
        __syncthreads();
        if (active) {
            const int tw_step = kTwiddleSize / (p * R);
#pragma unroll
            for (int bi = 0; bi < 8; ++bi) {
                if (bi < nb) {
                    const int i = t + bi * eighth;
                    const int k = i & (p - 1);
                    float2* ub = u + bi * R;
                    twiddle_and_dft<INV>(ub, lr, k, table, tw_step);
                    const int j = (i - k) * R + k;
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        if (r < R) {
                            buf[lds_pad(j + r * p)] = ub[r];
                        }
                    }
                }
            }
        }
        __syncthreads();
        p *= R;
    }
}

}  // namespace rf
