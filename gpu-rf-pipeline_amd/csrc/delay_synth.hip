// Synthetic capture pairs for the learned delay search: the same generator makes the training
// set and the evaluation set (different seeds), so both follow one stated distribution. It has
// no counterpart in the reference, whose fixtures are written by tools/gen_fixtures.py on the
// host; the delayed-channel construction follows that script's ch0/ch1 pair
// (tools/gen_fixtures.py:131-233: b is a delayed by d, so the GCC-PHAT peak sits at lag +d)
// with a fractional delay and coloured noise added.
//
// Pair p (global index) of seed s, n samples per channel, every random word from
// Philox4x32-10 with key (s mod 2^32, s / 2^32) and counter (index, p, stream, 0):
//   parameters (stream 0, index 0): words x0..x3, U(x) = (x >> 8) 2^-24,
//     d = -512 + 1024 U(x0)  (true delay, samples; exact in float32),
//     snr_db = -10 + 30 U(x1),  rho_a = 0.9 U(x2),  rho_b = 0.9 U(x3);
//   source s[m], m < n + 1152 (stream 1, index m): CN(0, 1) by Box-Muller from words x0, x1;
//   noise e_c[i] (stream 2 for a, 3 for b, index i): CN(0, 1) likewise; unit-power AR(1)
//     w_c[0] = e_c[0], w_c[i] = rho_c w_c[i-1] + sqrt(1 - rho_c^2) e_c[i];
//   a[i] = s[i + 576] + sigma w_a[i],
//   b[i] = sum_{j=-32..32} h[j] s[i + 576 - floor(d) - j] + sigma w_b[i],
//     h[j] = sinc(j - f) kaiser(j) / sum_j(sinc(j - f) kaiser(j)), f = d - floor(d),
//     kaiser(j) = I0(8 sqrt(1 - (j/32)^2)) / I0(8)  (a 65-tap windowed-sinc fractional delay),
//   sigma = 10^(-snr_db / 20).

#include "philox.hpp"
#include "fir_internal.hpp"
#include "rf_common.hpp"

#include <algorithm>
#include <cstdint>

namespace rf {
namespace {

constexpr int kSourceOffset = 576;
constexpr int kSourceExtra = 1152;
constexpr int kSynthThreads = 256;

struct PairParams {
    float delay;
    float sigma;
    float rho[2];
};

__device__ __forceinline__ PairParams pair_params(uint32_t k0, uint32_t k1, uint32_t pair)
{
    const Philox4 x = philox4x32_10(Philox4{{0u, pair, 0u, 0u}}, k0, k1);
    PairParams q;
    q.delay = -512.0f + 1024.0f * uniform24(x.v[0]);
    const float snr_db = -10.0f + 30.0f * uniform24(x.v[1]);
    q.sigma = exp10f(-snr_db / 20.0f);
    q.rho[0] = 0.9f * uniform24(x.v[2]);
    q.rho[1] = 0.9f * uniform24(x.v[3]);
    return q;
}

__device__ __forceinline__ float2 stream_normal(uint32_t k0, uint32_t k1, uint32_t index,
                                                uint32_t pair, uint32_t stream)
{
    const Philox4 x = philox4x32_10(Philox4{{index, pair, stream, 0u}}, k0, k1);
    return complex_normal(x.v[0], x.v[1]);
}

__global__ void synth_source_kernel(uint32_t k0, uint32_t k1, int first, int count, int span,
                                    float2* __restrict__ src)
{
    const long long total = static_cast<long long>(count) * span;
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < total;
         e += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int k = static_cast<int>(e / span);
        const int m = static_cast<int>(e - static_cast<long long>(k) * span);
        src[e] = stream_normal(k0, k1, static_cast<uint32_t>(m), static_cast<uint32_t>(first + k),
                               1u);
    }
}

// Modified Bessel function I0 by its power series sum_k ((x/2)^(2k)) / (k!)^2 (x <= 8 here:
// 40 terms leave a relative remainder below 1e-20).
__device__ double bessel_i0(double x)
{
    const double q = 0.25 * x * x;
    double term = 1.0;
    double sum = 1.0;
    for (int k = 1; k < 40; ++k) {
        term *= q / (static_cast<double>(k) * k);
        sum += term;
    }
    return sum;
}

// One workgroup per (pair, channel): the AR(1) noise as a two-pass chunked scan (each lane
// owns a contiguous chunk; pass 1 finds each chunk's zero-state end value, lane 0 chains the
// chunk carries, pass 2 reruns each chunk from its carry and writes signal + noise).
__global__ __launch_bounds__(kSynthThreads) void synth_channel_kernel(
    uint32_t k0, uint32_t k1, int first, int n, int span, const float2* __restrict__ src,
    float2* __restrict__ a, float2* __restrict__ b, float* __restrict__ delays)
{
    __shared__ double hd[65];
    __shared__ float h[65];
    __shared__ float2 carry[kSynthThreads + 1];
    __shared__ float decay[kSynthThreads];
    const int k = blockIdx.x;
    const int ch = blockIdx.y;
    const int t = threadIdx.x;
    const uint32_t pair = static_cast<uint32_t>(first + k);
    const PairParams q = pair_params(k0, k1, pair);
    const float delay = q.delay;
    const int di = static_cast<int>(floorf(delay));
    if (ch == 1) {
        if (t < 65) {
            const double f = static_cast<double>(delay) - di;
            const double x = static_cast<double>(t - 32) - f;
            const double px = 3.14159265358979323846 * x;
            const double sinc = x == 0.0 ? 1.0 : sin(px) / px;
            const double r = static_cast<double>(t - 32) / 32.0;
            hd[t] = sinc * bessel_i0(8.0 * sqrt(1.0 - r * r)) / bessel_i0(8.0);
        }
        __syncthreads();
        if (t == 0) {
            double sum = 0.0;
            for (int j = 0; j < 65; ++j) {
                sum += hd[j];
            }
            for (int j = 0; j < 65; ++j) {
                h[j] = static_cast<float>(hd[j] / sum);
            }
        }
    } else if (t == 0 && delays != nullptr) {
        delays[k] = delay;
    }
    const float rho = q.rho[ch];
    const float innov = sqrtf(1.0f - rho * rho);
    const uint32_t stream = 2u + static_cast<uint32_t>(ch);
    const int chunk = (n + kSynthThreads - 1) / kSynthThreads;
    const int start = t * chunk < n ? t * chunk : n;
    const int stop = start + chunk < n ? start + chunk : n;
    float2 z = make_float2(0.0f, 0.0f);
    for (int i = start; i < stop; ++i) {
        const float2 e = stream_normal(k0, k1, static_cast<uint32_t>(i), pair, stream);
        z = i == 0 ? e : make_float2(rho * z.x + innov * e.x, rho * z.y + innov * e.y);
    }
    carry[t + 1] = z;
    decay[t] = powf(rho, static_cast<float>(stop - start));
    __syncthreads();
    if (t == 0) {
        carry[0] = make_float2(0.0f, 0.0f);
        for (int c = 1; c <= kSynthThreads; ++c) {  // carry[c] = state after chunk c - 1
            const float2 end = carry[c];
            carry[c] = make_float2(decay[c - 1] * carry[c - 1].x + end.x,
                                   decay[c - 1] * carry[c - 1].y + end.y);
        }
    }
    __syncthreads();
    z = carry[t];
    const float2* s = src + static_cast<long long>(k) * span + kSourceOffset;
    float2* out = (ch == 0 ? a : b) + static_cast<long long>(k) * n;
    for (int i = start; i < stop; ++i) {
        const float2 e = stream_normal(k0, k1, static_cast<uint32_t>(i), pair, stream);
        z = i == 0 ? e : make_float2(rho * z.x + innov * e.x, rho * z.y + innov * e.y);
        float2 sig;
        if (ch == 0) {
            sig = s[i];
        } else {
            sig = make_float2(0.0f, 0.0f);
            const float2* base = s + i - di + 32;  // s[i - di - j] for j = -32 at base[0]
            for (int j = 0; j < 65; ++j) {
                const float2 v = base[-j];
                sig.x += h[j] * v.x;
                sig.y += h[j] * v.y;
            }
        }
        out[i] = make_float2(sig.x + q.sigma * z.x, sig.y + q.sigma * z.y);
    }
}

}  // namespace

void synth_pairs(uint64_t seed, int first, int count, int n, float2* a, float2* b,
                 float* delays, hipStream_t stream)
{
    const uint32_t k0 = static_cast<uint32_t>(seed);
    const uint32_t k1 = static_cast<uint32_t>(seed >> 32);
    const int span = n + kSourceExtra;
    StreamScratch src(static_cast<std::size_t>(count) * span * sizeof(float2), stream);
    const long long total = static_cast<long long>(count) * span;
    const unsigned grid = static_cast<unsigned>(std::min<long long>(ceil_div(total, 256),
                                                                    8LL * cu_count()));
    synth_source_kernel<<<grid, 256, 0, stream>>>(k0, k1, first, count, span,
                                                  static_cast<float2*>(src.ptr));
    check_launch("synth_source_kernel");
    synth_channel_kernel<<<dim3(static_cast<unsigned>(count), 2), kSynthThreads, 0, stream>>>(
        k0, k1, first, n, span, static_cast<const float2*>(src.ptr), a, b, delays);
    check_launch("synth_channel_kernel");
}

}  // namespace rf

extern "C" rf_status rf_synth_pairs(unsigned long long seed, int first_pair, int count, int n,
                                    rf_cf32* a, rf_cf32* b, float* delays, rf_stream stream)
{
    return rf::guarded([&] {
        if (a == nullptr || b == nullptr) rf::fail(RF_INVALID_ARG, "synth pairs: NULL output");
        if (count < 1 || n < 1 || first_pair < 0) {
            rf::fail(RF_INVALID_ARG, "synth pairs: need count >= 1, n >= 1, first_pair >= 0");
        }
        if (n > (1 << 24) || static_cast<long long>(first_pair) + count > (1LL << 31) - 1) {
            rf::fail(RF_LIMIT, "synth pairs: n above 2^24 or pair index above 2^31");
        }
        rf::synth_pairs(seed, first_pair, count, n, reinterpret_cast<float2*>(a),
                        reinterpret_cast<float2*>(b), delays, rf::as_stream(stream));
    });
}
