// Cross-correlation, GCC-PHAT and peak search for gfx950.
//
// Definition (src/xcorr.cu:1-7, tools/reference.py:19-22,114-125):
//   r[i] = sum_n b[n + k] conj(a[n]),  k = i - (num_a - 1),  i < num_a + num_b - 1.
// FFT form (src/xcorr.cu:235-375): with L = next power of two >= num_a + num_b - 1 and a
// stored rotated left by num_a - 1 (a[num_a-1] at 0, a[0 .. num_a-2] at the end), the circular
// correlation IFFT(FFT(b) conj(FFT(a))) / L holds lag i - (num_a - 1) at index i. GCC-PHAT
// (src/xcorr.cu:219-233) scales each product to unit magnitude, 0 where it is 0.
//
// MI355X layout: the rotation and the zero padding are index arithmetic in the loads, not
// copies; for L <= 4096 one workgroup does both forward transforms, the weighting, the
// inverse transform and the store of r; for longer L the four-step passes keep the spectrum
// in transposed order and the row workgroup fuses forward row FFTs of a and b, the weighting
// and the inverse row FFT, so a spectrum is never written to HBM in natural order.

#include "fft_plan.hpp"
#include "twiddle.hpp"
#include "xcorr_common.hpp"

#include <cstring>
#include <string>

struct rf_xcorr_fft {
    int num_a = 0;
    int num_b = 0;
    int phat = 0;
    int log2L = 0;
    rf::FourStep fs{};
    float2* tables = nullptr;
    float2* work = nullptr;  // 2 L values for L > 4096
    std::size_t bytes = 0;
};

namespace rf {
namespace {

// --- direct (time-domain) variants --------------------------------------------------------

// One lane per lag, both inputs from global memory (src/xcorr.cu:22-48).
__global__ void xcorr_naive_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                   float2* __restrict__ r, int num_a, int num_b)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<long long>(num_a) + num_b - 1) {
        return;
    }
    const long long lag = i - (num_a - 1);
    const long long first = lag < 0 ? -lag : 0;
    const long long end = (num_b - lag < num_a) ? num_b - lag : num_a;
    float re = 0.0f;
    float im = 0.0f;
    for (long long n = first; n < end; ++n) {
        const float2 sa = a[n];
        const float2 sb = b[n + lag];
        re = fmaf(sb.x, sa.x, re);
        re = fmaf(sb.y, sa.y, re);
        im = fmaf(sb.y, sa.x, im);
        im = fmaf(-sb.x, sa.y, im);
    }
    r[i] = make_float2(re, im);
}

// Chunks of a (blockDim.x samples) and b (2 blockDim.x - 1) staged in LDS, as
// src/xcorr.cu:99-156. Dynamic LDS: (3 blockDim.x - 1) values.
__global__ void xcorr_tiled_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                   float2* __restrict__ r, int num_a, int num_b)
{
    extern __shared__ float2 tile[];
    float2* tile_a = tile;
    float2* tile_b = tile + blockDim.x;
    const int chunk = blockDim.x;
    const int tid = threadIdx.x;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + tid;
    const bool active = i < static_cast<long long>(num_a) + num_b - 1;
    const long long lag0 = static_cast<long long>(blockIdx.x) * blockDim.x - (num_a - 1);
    const long long last_lag = (lag0 + chunk - 1 < num_b - 1) ? lag0 + chunk - 1 : num_b - 1;
    const long long n_first = last_lag < 0 ? -last_lag : 0;
    const long long n_end = (num_b - lag0 < num_a) ? num_b - lag0 : num_a;
    float re = 0.0f;
    float im = 0.0f;
    for (long long n0 = n_first; n0 < n_end; n0 += chunk) {
        const long long na = n0 + tid;
        tile_a[tid] = na < num_a ? a[na] : make_float2(0.0f, 0.0f);
        for (int j = tid; j < 2 * chunk - 1; j += chunk) {
            const long long m = n0 + lag0 + j;
            tile_b[j] = (m >= 0 && m < num_b) ? b[m] : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        if (active) {
            for (int j = 0; j < chunk; ++j) {
                const float2 sa = tile_a[j];
                const float2 sb = tile_b[j + tid];
                re = fmaf(sb.x, sa.x, re);
                re = fmaf(sb.y, sa.y, re);
                im = fmaf(sb.y, sa.x, im);
                im = fmaf(-sb.x, sa.y, im);
            }
        }
        __syncthreads();
    }
    if (active) {
        r[i] = make_float2(re, im);
    }
}

// --- FFT form -----------------------------------------------------------------------------

// L <= 4096: the whole correlation in one workgroup of max(L/8, 1) lanes... L >= 8 here.
__global__ __launch_bounds__(512) void xcorr_fft_small_kernel(const float2* __restrict__ a,
                                                              const float2* __restrict__ b,
                                                              float2* __restrict__ r,
                                                              int num_a, int num_b, int log2L,
                                                              int phat,
                                                              const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const int L = 1 << log2L;
    const int slots = lds_slots(L);
    float2* fa = lds;
    float2* fb = lds + slots;
    for (int n = threadIdx.x; n < L; n += blockDim.x) {
        fa[lds_pad(n)] = rotated_a(a, num_a, L, n);
        fb[lds_pad(n)] = padded_b(b, num_b, n);
    }
    __syncthreads();
    lds_fft<false>(fa, log2L, threadIdx.x, true, table);
    lds_fft<false>(fb, log2L, threadIdx.x, true, table);
    const float scale = 1.0f / static_cast<float>(L);
    for (int n = threadIdx.x; n < L; n += blockDim.x) {
        fb[lds_pad(n)] = weigh(fa[lds_pad(n)], fb[lds_pad(n)], scale, phat);
    }
    __syncthreads();
    lds_fft<true>(fb, log2L, threadIdx.x, true, table);
    const long long lags = static_cast<long long>(num_a) + num_b - 1;
    for (int n = threadIdx.x; n < lags; n += blockDim.x) {
        r[n] = fb[lds_pad(n)];
    }
}

// L in {1, 2, 4}: one lane, direct DFTs.
__global__ void xcorr_fft_tiny_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                      float2* __restrict__ r, int num_a, int num_b, int L,
                                      int phat)
{
    float2 va[4];
    float2 vb[4];
    for (int n = 0; n < L; ++n) {
        va[n] = rotated_a(a, num_a, L, n);
        vb[n] = padded_b(b, num_b, n);
    }
    float2 fa[4];
    float2 fb[4];
    // cos/sin of multiples of pi/2 are exact
    for (int k = 0; k < L; ++k) {
        fa[k] = make_float2(0.0f, 0.0f);
        fb[k] = make_float2(0.0f, 0.0f);
        for (int n = 0; n < L; ++n) {
            const int e = (n * k) % L;  // exp(-2 pi i e / L)
            const int q = e * (4 / L);  // quarter turns
            const float2 w = q == 0 ? make_float2(1, 0)
                             : q == 1 ? make_float2(0, -1)
                             : q == 2 ? make_float2(-1, 0)
                                      : make_float2(0, 1);
            fa[k] = cadd(fa[k], cmul(va[n], w));
            fb[k] = cadd(fb[k], cmul(vb[n], w));
        }
    }
    const float scale = 1.0f / static_cast<float>(L);
    float2 wgt[4];
    for (int k = 0; k < L; ++k) {
        wgt[k] = weigh(fa[k], fb[k], scale, phat);
    }
    const int lags = num_a + num_b - 1;
    for (int n = 0; n < lags; ++n) {
        float2 acc = make_float2(0.0f, 0.0f);
        for (int k = 0; k < L; ++k) {
            const int e = (n * k) % L;
            const int q = e * (4 / L);
            const float2 w = q == 0 ? make_float2(1, 0)
                             : q == 1 ? make_float2(0, 1)
                             : q == 2 ? make_float2(-1, 0)
                                      : make_float2(0, -1);
            acc = cadd(acc, cmul(wgt[k], w));
        }
        r[n] = acc;
    }
}

// Column pass of a (blockIdx.y == 0) or b (== 1) straight from the caller's buffers.
__global__ __launch_bounds__(512) void xcorr_column_fwd_kernel(
    const float2* __restrict__ a, const float2* __restrict__ b, float2* __restrict__ work,
    int num_a, int num_b, FourStep fs, const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const long long L = 1LL << fs.log2L;
    const int log2n1 = fs.log2n1;
    const long long c0 = static_cast<long long>(blockIdx.x) * column_tile(fs);
    float2* dst = work + (blockIdx.y ? L : 0);
    auto store = [=](long long j1, int k2, float2 v) {
        dst[j1 + (static_cast<long long>(k2) << log2n1)] = v;
    };
    if (blockIdx.y == 0) {
        column_pass_block<false>(
            fs, c0,
            [=](long long j1, int j2) {
                return rotated_a(a, num_a, L, j1 + (static_cast<long long>(j2) << log2n1));
            },
            store, lds, table);
    } else {
        column_pass_block<false>(
            fs, c0,
            [=](long long j1, int j2) {
                return padded_b(b, num_b, j1 + (static_cast<long long>(j2) << log2n1));
            },
            store, lds, table);
    }
}

// Row k2 of both spectra: forward row FFTs, weighting, inverse row FFT, conjugate twiddle;
// the result replaces row k2 of b's work buffer.
__global__ __launch_bounds__(512) void xcorr_row_kernel(float2* __restrict__ work, FourStep fs,
                                                        int phat,
                                                        const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const long long L = 1LL << fs.log2L;
    const int n1 = 1 << fs.log2n1;
    const int n2 = 1 << fs.log2n2;
    const int F = lds_batch_per_block(fs.log2n1);
    const int slots = lds_slots(n1);
    const int row0 = blockIdx.x * F;
    const int count = n2 - row0 < F ? n2 - row0 : F;
    float2* wa = work + static_cast<long long>(row0) * n1;
    float2* wb = work + L + static_cast<long long>(row0) * n1;
    float2* la = lds;
    float2* lb = lds + F * slots;
    const int total = count * n1;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int s = (e >> fs.log2n1) * slots + lds_pad(e & (n1 - 1));
        la[s] = wa[e];
        lb[s] = wb[e];
    }
    __syncthreads();
    const int per = n1 >> 3;
    const int f = threadIdx.x / per;
    const int t = threadIdx.x - f * per;
    const int fo = (f < count ? f : 0) * slots;
    lds_fft<false>(la + fo, fs.log2n1, t, f < count, table);
    lds_fft<false>(lb + fo, fs.log2n1, t, f < count, table);
    const float scale = 1.0f / static_cast<float>(L);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int s = (e >> fs.log2n1) * slots + lds_pad(e & (n1 - 1));
        lb[s] = weigh(la[s], lb[s], scale, phat);
    }
    __syncthreads();
    lds_fft<true>(lb + fo, fs.log2n1, t, f < count, table);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int rr = e >> fs.log2n1;
        const int j1 = e & (n1 - 1);
        const float2 v = lb[rr * slots + lds_pad(j1)];
        wb[e] = cmul(v, four_step_twiddle<true>(fs, static_cast<long long>(j1) * (row0 + rr)));
    }
}

// Row pass for 4096-point rows: one row per workgroup, 256 lanes, register-resident radix-16
// transforms (same arithmetic as xcorr_row_kernel up to the transform's rounding).
__global__ __launch_bounds__(k4096Lanes) void xcorr_row4096_kernel(float2* __restrict__ work,
                                                                   FourStep fs, int phat,
                                                                   const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    const long long L = 1LL << fs.log2L;
    const long long row = blockIdx.x;
    xcorr_row4096(work + (row << k4096Log2), work + L + (row << k4096Log2), fs, row, phat,
                  1.0f / static_cast<float>(L), buf, stages);
}

// Inverse column pass of b's work buffer into r (natural order, first num_a + num_b - 1).
__global__ __launch_bounds__(512) void xcorr_column_inv_kernel(const float2* __restrict__ work,
                                                               float2* __restrict__ r,
                                                               long long lags, FourStep fs,
                                                               const float2* __restrict__ table)
{
    // The conjugate twiddles were applied by the row kernel, so this pass is a plain column
    // FFT (unlike column_pass_block, which twiddles after the transform), run from registers
    // to registers as there (stockham_column): the launch has C n2 / 8 lanes, lane tid takes
    // column tid mod C and rows t + (n2 / 8) r, and stores its outputs itself.
    extern __shared__ float2 lds[];
    const long long L = 1LL << fs.log2L;
    const int log2n1 = fs.log2n1;
    const float2* src = work + L;
    const long long c0 = static_cast<long long>(blockIdx.x) * column_tile(fs);
    const int n2 = 1 << fs.log2n2;
    const int n8 = n2 >> 3;
    const int lc = column_tile_log2(fs);
    const int C = 1 << lc;
    const int slots = column_slots(n2);
    const int col = threadIdx.x & (C - 1);
    const int t = threadIdx.x >> lc;
    const long long j1 = c0 + col;
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        v[q] = src[j1 + (static_cast<long long>(t + q * n8) << log2n1)];
    }
    const int lr = stockham_column<true>(lds + col * slots, v, fs.log2n2, t, true, table);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int j2 = t + stockham_order(lr, q) * n8;
        const long long n = j1 + (static_cast<long long>(j2) << log2n1);
        if (n < lags) {
            r[n] = v[q];
        }
    }
}

// --- peak search (src/main.cpp:390-401 moved to the device) -------------------------------

__global__ void peak_init_kernel(rf_peak_result* result)
{
    *reinterpret_cast<unsigned long long*>(result) = 0ULL;
    result->frac = 0.0;
}

__global__ __launch_bounds__(256) void peak_reduce_kernel(const float2* __restrict__ r, int n,
                                                          rf_peak_result* result)
{
    unsigned long long best = 0ULL;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float2 v = r[i];
        const unsigned long long key = peak_key(hypotf(v.x, v.y), i);
        best = key > best ? key : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long other = __shfl_xor(best, off, 64);
        best = other > best ? other : best;
    }
    __shared__ unsigned long long wave_best[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        wave_best[wave] = best;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < static_cast<int>(blockDim.x >> 6); ++w) {
            best = wave_best[w] > best ? wave_best[w] : best;
        }
        atomicMax(reinterpret_cast<unsigned long long*>(result), best);
    }
}

// Turns the key into (index, magnitude) and fits a parabola through |r| at index - 1, index,
// index + 1 in float64: frac = (m- - m+) / (2 (m- - 2 m0 + m+)), clamped to [-0.5, 0.5].
__global__ void peak_finish_kernel(const float2* __restrict__ r, int n, rf_peak_result* result)
{
    const unsigned long long key = *reinterpret_cast<const unsigned long long*>(result);
    const int idx = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(key & 0xFFFFFFFFULL));
    const float mag = __uint_as_float(static_cast<unsigned>(key >> 32));
    double frac = 0.0;
    if (idx > 0 && idx + 1 < n) {
        const float2 lo = r[idx - 1];
        const float2 hi = r[idx + 1];
        const double m0 = static_cast<double>(mag);
        const double ml = hypot(static_cast<double>(lo.x), static_cast<double>(lo.y));
        const double mh = hypot(static_cast<double>(hi.x), static_cast<double>(hi.y));
        const double den = ml - 2.0 * m0 + mh;
        if (den < 0.0) {
            frac = 0.5 * (ml - mh) / den;
            frac = frac > 0.5 ? 0.5 : (frac < -0.5 ? -0.5 : frac);
        }
    }
    result->index = idx;
    result->magnitude = mag;
    result->frac = frac;
}

const char* const kXcorrNames[RF_XCORR_IMPL_COUNT] = {"naive", "tiled", "fft"};
const char* const kXcorrSummaries[RF_XCORR_IMPL_COUNT] = {
    "one lane per lag, global memory only",
    "both inputs staged in LDS tiles, chunk by chunk",
    "the library's FFT: forward FFTs, conjugate multiply, inverse FFT",
};

void check_lengths(int num_a, int num_b, const char* what)
{
    if (num_a < 1 || num_b < 1) {
        fail(RF_INVALID_ARG, std::string(what) + ": both inputs need at least one sample");
    }
    if (static_cast<long long>(num_a) + num_b - 1 > 0x7FFFFFFFLL) {
        fail(RF_LIMIT, std::string(what) + ": more lags than fit in an int");
    }
}

}  // namespace

void xcorr_fft_run(const rf_xcorr_fft& h, const float2* a, const float2* b, float2* r,
                   hipStream_t stream)
{
    const float2* table = twiddle_table_4096();
    const int L = 1 << h.log2L;
    if (h.log2L < 3) {
        xcorr_fft_tiny_kernel<<<1, 1, 0, stream>>>(a, b, r, h.num_a, h.num_b, L, h.phat);
        check_launch("xcorr_fft_tiny_kernel");
        return;
    }
    if (h.log2L <= kTwiddleLog2) {
        const std::size_t bytes = 2 * static_cast<std::size_t>(lds_slots(L)) * sizeof(float2);
        xcorr_fft_small_kernel<<<1, L / 8, bytes, stream>>>(a, b, r, h.num_a, h.num_b, h.log2L,
                                                           h.phat, table);
        check_launch("xcorr_fft_small_kernel");
        return;
    }
    const FourStep& fs = h.fs;
    const int n2 = 1 << fs.log2n2;
    const int C = column_tile(fs);
    const std::size_t cbytes = static_cast<std::size_t>(C) * column_slots(n2) * sizeof(float2);
    const dim3 cgrid(static_cast<unsigned>((1 << fs.log2n1) / C), 2);
    xcorr_column_fwd_kernel<<<cgrid, C * n2 / 8, cbytes, stream>>>(a, b, h.work, h.num_a,
                                                                   h.num_b, fs, table);
    check_launch("xcorr_column_fwd_kernel");
    if (fs.log2n1 == k4096Log2) {
        xcorr_row4096_kernel<<<static_cast<unsigned>(n2), k4096Lanes, 0, stream>>>(
            h.work, fs, h.phat, fft4096_stage_table());
        check_launch("xcorr_row4096_kernel");
    } else {
        const int F = lds_batch_per_block(fs.log2n1);
        const std::size_t rbytes =
            2 * static_cast<std::size_t>(F) * lds_slots(1 << fs.log2n1) * sizeof(float2);
        xcorr_row_kernel<<<static_cast<unsigned>(ceil_div(n2, F)), F * ((1 << fs.log2n1) >> 3),
                           rbytes, stream>>>(h.work, fs, h.phat, table);
        check_launch("xcorr_row_kernel");
    }
    const dim3 igrid(static_cast<unsigned>((1 << fs.log2n1) / C), 1);
    xcorr_column_inv_kernel<<<igrid, C * n2 / 8, cbytes, stream>>>(
        h.work, r, static_cast<long long>(h.num_a) + h.num_b - 1, fs, table);
    check_launch("xcorr_column_inv_kernel");
}

void peak_launch(const float2* r, int n, rf_peak_result* result, hipStream_t stream)
{
    peak_init_kernel<<<1, 1, 0, stream>>>(result);
    check_launch("peak_init_kernel");
    const long long want = ceil_div(n, 256);
    const unsigned grid = static_cast<unsigned>(want < 1024 ? (want < 1 ? 1 : want) : 1024);
    peak_reduce_kernel<<<grid, 256, 0, stream>>>(r, n, result);
    check_launch("peak_reduce_kernel");
    peak_finish_kernel<<<1, 1, 0, stream>>>(r, n, result);
    check_launch("peak_finish_kernel");
}

}  // namespace rf

using rf::guarded;

namespace {

void destroy_xcorr(rf_xcorr_fft* h)
{
    if (h != nullptr) {
        (void)hipFree(h->tables);
        (void)hipFree(h->work);
        delete h;
    }
}

}  // namespace

extern "C" {

int rf_xcorr_impl_count(void)
{
    return RF_XCORR_IMPL_COUNT;
}

const char* rf_xcorr_impl_name(int impl)
{
    return (impl >= 0 && impl < RF_XCORR_IMPL_COUNT) ? rf::kXcorrNames[impl] : nullptr;
}

const char* rf_xcorr_impl_summary(int impl)
{
    return (impl >= 0 && impl < RF_XCORR_IMPL_COUNT) ? rf::kXcorrSummaries[impl] : nullptr;
}

rf_status rf_xcorr_parse_impl(const char* name, int* impl)
{
    return guarded([&] {
        if (name == nullptr || impl == nullptr) {
            rf::fail(RF_INVALID_ARG, "xcorr: NULL implementation name");
        }
        for (int i = 0; i < RF_XCORR_IMPL_COUNT; ++i) {
            if (std::strcmp(name, rf::kXcorrNames[i]) == 0) {
                *impl = i;
                return;
            }
        }
        rf::fail(RF_INVALID_ARG, std::string("unknown xcorr implementation '") + name + "'");
    });
}

rf_status rf_xcorr_direct(int impl, const rf_cf32* a, const rf_cf32* b, rf_cf32* r, int num_a,
                          int num_b, rf_stream stream)
{
    return guarded([&] {
        rf::check_lengths(num_a, num_b, "xcorr");
        if (a == nullptr || b == nullptr || r == nullptr) {
            rf::fail(RF_INVALID_ARG, "xcorr: NULL buffer");
        }
        const long long lags = static_cast<long long>(num_a) + num_b - 1;
        const int block = 256;
        const unsigned grid = static_cast<unsigned>(rf::ceil_div(lags, block));
        hipStream_t s = rf::as_stream(stream);
        const float2* fa = reinterpret_cast<const float2*>(a);
        const float2* fb = reinterpret_cast<const float2*>(b);
        float2* fr = reinterpret_cast<float2*>(r);
        if (impl == RF_XCORR_NAIVE) {
            rf::xcorr_naive_kernel<<<grid, block, 0, s>>>(fa, fb, fr, num_a, num_b);
            rf::check_launch("xcorr_naive_kernel");
        } else if (impl == RF_XCORR_TILED) {
            const std::size_t bytes = (3 * static_cast<std::size_t>(block) - 1) * sizeof(float2);
            rf::xcorr_tiled_kernel<<<grid, block, bytes, s>>>(fa, fb, fr, num_a, num_b);
            rf::check_launch("xcorr_tiled_kernel");
        } else {
            rf::fail(RF_INVALID_ARG, "xcorr direct: impl must be naive or tiled");
        }
    });
}

rf_status rf_xcorr_fft_create(int num_a, int num_b, int phat, rf_xcorr_fft** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "xcorr fft: NULL output handle");
        *out = nullptr;
        rf::check_lengths(num_a, num_b, "xcorr fft");
        const long long lags = static_cast<long long>(num_a) + num_b - 1;
        long long L = 1;
        while (L < lags) {
            L *= 2;
        }
        if (L > (1LL << 24)) {
            rf::fail(RF_LIMIT, "xcorr fft: " + std::to_string(lags)
                                   + " lags need an FFT longer than the supported 2^24 points");
        }
        rf_xcorr_fft* h = new rf_xcorr_fft;
        h->num_a = num_a;
        h->num_b = num_b;
        h->phat = phat ? 1 : 0;
        h->log2L = rf::ilog2(L);
        try {
            (void)rf::twiddle_table_4096();
            (void)rf::fft4096_stage_table();
            if (h->log2L > rf::kTwiddleLog2) {
                h->fs = rf::make_four_step_shape(h->log2L);
                const long long tsize = rf::four_step_table_size(h->log2L);
                RF_HIP_CHECK(hipMalloc(&h->tables, tsize * sizeof(float2)));
                rf::upload_four_step_tables(h->log2L, h->tables);
                h->fs.lo = h->tables;
                h->fs.hi = h->tables + rf::kTwiddleSize;
                RF_HIP_CHECK(hipMalloc(&h->work, 2 * L * sizeof(float2)));
                h->bytes = (tsize + 2 * L) * sizeof(float2);
            }
        } catch (...) {
            destroy_xcorr(h);
            throw;
        }
        *out = h;
    });
}

rf_status rf_xcorr_fft_destroy(rf_xcorr_fft* handle)
{
    return guarded([&] { destroy_xcorr(handle); });
}

rf_status rf_xcorr_fft_bytes(const rf_xcorr_fft* handle, size_t* bytes)
{
    return guarded([&] {
        if (handle == nullptr || bytes == nullptr) rf::fail(RF_INVALID_ARG, "xcorr fft: NULL");
        *bytes = handle->bytes;
    });
}

rf_status rf_xcorr_fft_length(const rf_xcorr_fft* handle, int* fft_length)
{
    return guarded([&] {
        if (handle == nullptr || fft_length == nullptr) rf::fail(RF_INVALID_ARG, "xcorr fft: NULL");
        *fft_length = 1 << handle->log2L;
    });
}

rf_status rf_xcorr_fft_run(const rf_xcorr_fft* handle, const rf_cf32* a, const rf_cf32* b,
                           rf_cf32* r, rf_stream stream)
{
    return guarded([&] {
        if (handle == nullptr) rf::fail(RF_INVALID_ARG, "xcorr fft: NULL handle");
        if (a == nullptr || b == nullptr || r == nullptr) {
            rf::fail(RF_INVALID_ARG, "xcorr fft: NULL buffer");
        }
        rf::xcorr_fft_run(*handle, reinterpret_cast<const float2*>(a),
                          reinterpret_cast<const float2*>(b), reinterpret_cast<float2*>(r),
                          rf::as_stream(stream));
    });
}

rf_status rf_peak(const rf_cf32* r, int n, rf_peak_result* result, rf_stream stream)
{
    return guarded([&] {
        if (n < 1) rf::fail(RF_INVALID_ARG, "peak: need at least one value");
        if (r == nullptr || result == nullptr) rf::fail(RF_INVALID_ARG, "peak: NULL buffer");
        rf::peak_launch(reinterpret_cast<const float2*>(r), n, result, rf::as_stream(stream));
    });
}

}  // extern "C"
