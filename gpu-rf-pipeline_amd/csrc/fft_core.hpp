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
#include "tables.hpp"

namespace rf {

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

// u *= exp(-+2 pi i e / 16) for a compile-time e (after full unrolling); the quarter turns
// are exact swaps.
template <bool INV>
__device__ __forceinline__ float2 rot16(float2 u, int e)
{
    e &= 15;
    if (e == 0) {
        return u;
    }
    if (e == 4) {
        return INV ? make_float2(-u.y, u.x) : make_float2(u.y, -u.x);
    }
    if (e == 8) {
        return make_float2(-u.x, -u.y);
    }
    if (e == 12) {
        return INV ? make_float2(u.y, -u.x) : make_float2(-u.y, u.x);
    }
    // cos and sin of 2 pi e / 16 for e = 0..3 (the rest follow by quarter turns)
    const float c1 = 0.92387953251128675613f;  // cos(pi / 8)
    const float s1 = 0.38268343236508977173f;  // sin(pi / 8)
    const float h = 0.70710678118654752440f;   // cos(pi / 4)
    const float cs[4] = {1.0f, c1, h, s1};
    const float sn[4] = {0.0f, s1, h, c1};
    const int q = e & 3;
    const int quarter = e >> 2;
    float c = cs[q];
    float sv = sn[q];
    // rotate (c, s) by quarter * 90 degrees: (c, s) -> (-s, c)
    for (int i = 0; i < quarter; ++i) {
        const float tmp = c;
        c = -sv;
        sv = tmp;
    }
    // forward: multiply by (c - i s); inverse: by (c + i s)
    const float si = INV ? sv : -sv;
    return make_float2(u.x * c - u.y * si, u.x * si + u.y * c);
}

// a <- a + w b, b <- a - w b in six instructions: the sum as two FMA chains per component and
// the difference as 2 a - sum (one FMA per component), instead of a complex multiply (four) and
// an add and a subtract (four). The difference carries the sum's rounding (about one ulp of
// the larger of the two results), as a plain butterfly after a rounded product does.
__device__ __forceinline__ void fused_bfly(float2& a, float2& b, float2 w)
{
    const float sx = fmaf(w.x, b.x, fmaf(-w.y, b.y, a.x));
    const float sy = fmaf(w.x, b.y, fmaf(w.y, b.x, a.y));
    b = make_float2(fmaf(2.0f, a.x, -sx), fmaf(2.0f, a.y, -sy));
    a = make_float2(sx, sy);
}

// a <- a + rot_e(c), c <- a - rot_e(c), rot_e = exp(-+2 pi i e / 16) for a compile-time e
// (after unrolling): quarter turns are exact swaps and a plain butterfly, the others
// fused_bfly with the constant factor.
template <bool INV>
__device__ __forceinline__ void rot_bfly(float2& a, float2& c, int e)
{
    if ((e & 3) == 0) {
        const float2 r = rot16<INV>(c, e);
        const float2 t = a;
        a = cadd(t, r);
        c = csub(t, r);
        return;
    }
    fused_bfly(a, c, rot16<INV>(make_float2(1.0f, 0.0f), e));
}

// Second stage of the 16-point DFT: y[4 n2 + k1] holds output k1 of the first-stage DFT4 of
// group n2, not yet multiplied by w16^(n2 k1); DFT4 over n2 with those factors folded into its
// butterflies, outputs X[k1 + 4 k2] into v in natural order.
template <bool INV>
__device__ __forceinline__ void dft16_second(const float2* y, float2* v)
{
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 a = y[k1];
        float2 b = rot16<INV>(y[4 + k1], k1);
        float2 c = y[8 + k1];
        float2 d = y[12 + k1];
        rot_bfly<INV>(a, c, 2 * k1);   // a + w^2k1 c, a - w^2k1 c
        rot_bfly<INV>(b, d, 3 * k1);   // w^k1 b + w^3k1 d, w^k1 b - w^3k1 d
        // forward: -i * d13 = (d13.y, -d13.x); inverse: +i * d13 = (-d13.y, d13.x)
        const float2 m = INV ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
        v[k1] = cadd(a, b);
        v[k1 + 8] = csub(a, b);
        v[k1 + 4] = cadd(c, m);
        v[k1 + 12] = csub(c, m);
    }
}

// 16-point DFT as 4 x 4: DFT4 over n1 of x[4 n1 + n2], twiddle w16^(n2 k1), DFT4 over n2;
// output X[k1 + 4 k2] back into v in natural order.
template <bool INV>
__device__ __forceinline__ void dft16(float2* v)
{
    float2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 a = v[n2];
        float2 b = v[4 + n2];
        float2 c = v[8 + n2];
        float2 d = v[12 + n2];
        dft4<INV>(a, b, c, d);
        y[4 * n2 + 0] = a;
        y[4 * n2 + 1] = b;
        y[4 * n2 + 2] = c;
        y[4 * n2 + 3] = d;
    }
    dft16_second<INV>(y, v);
}

// The 16-point DFT of (u[0], w[1] u[1], ..., w[15] u[15]) (conj(w[r]) for the inverse): in
// each first-stage DFT4 the factors of the third and fourth inputs are folded into its two
// input butterflies (fused_bfly), the first two are applied as products.
template <bool INV>
__device__ __forceinline__ void twiddled_dft16(float2* u, const float2* w)
{
    auto f = [&](int r) { return INV ? make_float2(w[r].x, -w[r].y) : w[r]; };
    float2 y[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        float2 p0 = n2 == 0 ? u[0] : cmul(u[n2], f(n2));
        float2 p1 = cmul(u[4 + n2], f(4 + n2));
        float2 p2 = u[8 + n2];
        float2 p3 = u[12 + n2];
        fused_bfly(p0, p2, f(8 + n2));    // s02, d02
        fused_bfly(p1, p3, f(12 + n2));   // s13, d13
        const float2 m = INV ? make_float2(-p3.y, p3.x) : make_float2(p3.y, -p3.x);
        y[4 * n2 + 0] = cadd(p0, p1);
        y[4 * n2 + 2] = csub(p0, p1);
        y[4 * n2 + 1] = cadd(p2, m);
        y[4 * n2 + 3] = csub(p2, m);
    }
    dft16_second<INV>(y, u);
}

// One radix-2^lr butterfly on u[0 .. R-1] with its Stockham twiddles. `k` = i mod p and
// `tw_step` = 4096 / (p R): u[r] *= W^(r k tw_step). Only W^m for m = base, 2 base and 4 base
// come from the table; W^3m = W^m W^2m, W^5m = W^m W^4m, W^6m = W^2m W^4m, W^7m = W^3m W^4m
// (at most two roundings more): 3 loads instead of 7 for a radix-8 butterfly, whose table
// loads otherwise outnumber its data accesses.
template <bool INV>
__device__ __forceinline__ void twiddle_and_dft(float2* u, int lr, int k,
                                                const float2* __restrict__ table, int tw_step)
{
    const int R = 1 << lr;
    if (k != 0) {
        const int base = k * tw_step;
        const float2 w1 = twiddle<INV>(table, base);
        u[1] = cmul(u[1], w1);
        if (R >= 4) {
            const float2 w2 = twiddle<INV>(table, 2 * base);
            const float2 w3 = cmul(w1, w2);
            u[2] = cmul(u[2], w2);
            u[3] = cmul(u[3], w3);
            if (R == 8) {
                const float2 w4 = twiddle<INV>(table, 4 * base);
                u[4] = cmul(u[4], w4);
                u[5] = cmul(u[5], cmul(w1, w4));
                u[6] = cmul(u[6], cmul(w2, w4));
                u[7] = cmul(u[7], cmul(w3, w4));
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

// Workgroup barrier that orders LDS accesses only: unlike __syncthreads() it does not wait for
// the lanes' outstanding global loads, so a load issued for later work stays in flight across
// it.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One pass of radix R = 2^LR (compile-time, so the eight values stay in registers): each
// thread reads eight values (one radix-8, two radix-4 or four radix-2 butterflies), waits at a
// barrier, and writes its results.
template <bool INV, int LR>
__device__ __forceinline__ void lds_pass(float2* buf, int P, int t, int p, bool active,
                                         const float2* __restrict__ table)
{
    constexpr int R = 1 << LR;
    constexpr int NB = 8 >> LR;  // butterflies per thread
    const int eighth = P >> 3;
    const int stride = P >> LR;  // P / R
    float2 u[8];
    if (active) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int bi = b >> LR;
            const int r = b & (R - 1);
            u[b] = buf[lds_pad(t + bi * eighth + r * stride)];
        }
    }
    __syncthreads();
    if (active) {
        const int tw_step = kTwiddleSize / (p * R);
#pragma unroll
        for (int bi = 0; bi < NB; ++bi) {
            const int i = t + bi * eighth;
            const int k = i & (p - 1);
            twiddle_and_dft<INV>(u + bi * R, LR, k, table, tw_step);
            const int j = (i - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                buf[lds_pad(j + r * p)] = u[bi * R + r];
            }
        }
    }
    __syncthreads();
}

// A transform of P = 2^log2P (8 <= P <= 4096) points done by P/8 threads of a workgroup on
// data already in LDS (padded layout), in place, natural order in and out. Every thread of the
// workgroup must call this (it has barriers); `t` is the thread's index within the transform
// and `active` is false for threads that only take part in the barriers.
template <bool INV>
__device__ __forceinline__ void lds_fft(float2* buf, int log2P, int t, bool active,
                        const float2* __restrict__ table)
{
    const int P = 1 << log2P;
    const int full8 = log2P / 3;
    int p = 1;
    for (int pass = 0; pass < full8; ++pass) {
        lds_pass<INV, 3>(buf, P, t, p, active, table);
        p <<= 3;
    }
    const int rem = log2P - 3 * full8;
    if (rem == 2) {
        lds_pass<INV, 2>(buf, P, t, p, active, table);
    } else if (rem == 1) {
        lds_pass<INV, 1>(buf, P, t, p, active, table);
    }
}

}  // namespace rf
