// Device side of the batched GCC-PHAT shared by rf_gcc_batch (gcc_batch.hip) and the learned
// delay search (delay_search.hip): per-pair descriptors, the forward column pass, the inverse
// column pass with the fused peak search and the sub-sample finish. Result records start with
// the 16 bytes of rf_peak_result (int32 index, float magnitude, double frac), whose first 8
// bytes double as the 64-bit peak key while the passes run.
#pragma once

#include "fft4096.hpp"
#include "fft_plan.hpp"
#include "xcorr_common.hpp"

#include <string>

namespace rf {

struct PairDesc {
    const float2* a;
    const float2* b;
    float2* r;        // optional full output (num_a + num_b - 1 values)
    float2* work;     // 2 L values
    const float2* lo;
    const float2* hi;
    int num_a;
    int num_b;
    int log2L;
    int log2n1;
    int log2n2;
    int col_items;    // n1 / C: inverse column-pass items (4096-value tiles)
    int fwd_items;    // forward column-pass items (8192-value tiles)
    int row_items;    // n2 / F
    int fused;        // 1: L = 2^15, one workgroup per pair (gcc_pair32k.hpp); no work buffer
};

namespace {

constexpr int kItemThreads = 512;

// The forward column pass works on tiles of 8192 values (C = 8192 / n2 columns, or all n1 if
// fewer) on 1024 lanes, twice the inverse pass's 4096, so that it reads and writes row segments
// twice as long (256 B per row at L = 2^20, 128 B at 2^21): measured 1.43 -> 1.04 ms per
// batch of 256 uneven pairs (2^12..2^20 samples). Its 80 KiB of LDS allow two workgroups per
// CU. The inverse pass, whose peak reduction ends every item with a workgroup-wide barrier,
// measured slower with the wide tile
// (0.59 -> 0.69 ms) and keeps 4096; with its transforms run from registers the wide tile
// (the waves' peaks then in the tile's LDS, so that two 80 KiB workgroups share a CU, and 64
// registers) ran that batch in the same 2.05 ms. Each pass has its own item map.
constexpr int kFwdTileLog2 = 13;
constexpr int kFwdThreads = 1 << (kFwdTileLog2 - 3);
constexpr int kFwdWorkgroupsPerCu = 2;

// Item maps: item -> {pair, index of the item within its pair}, built on the host, so a
// persistent workgroup finds its item's pair with one load.
__device__ __forceinline__ FourStep shape_of(const PairDesc& d)
{
    FourStep fs;
    fs.log2L = d.log2L;
    fs.log2n1 = d.log2n1;
    fs.log2n2 = d.log2n2;
    fs.lo = d.lo;
    fs.hi = d.hi;
    return fs;
}

template <class R>
__global__ void batch_init_kernel(R* results, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        *reinterpret_cast<unsigned long long*>(&results[i]) = 0ULL;
        results[i].frac = 0.0;
    }
}

// Forward column items begin .. begin + total - 1 of the forward map, for a and for b.
__global__ __launch_bounds__(kFwdThreads) void batch_column_fwd_kernel(
    const PairDesc* __restrict__ pairs, const int2* __restrict__ map, int begin, int total,
    const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    xcd_major_items(2 * total, [&](int item) {
        const int which = item >= total ? 1 : 0;
        const int2 m = map[begin + item - which * total];
        const PairDesc d = pairs[m.x];
        const FourStep fs = shape_of(d);
        // Indices below L <= 2^24 and lengths below 2^24: 32-bit arithmetic throughout.
        const int L = 1 << d.log2L;
        const int log2n1 = d.log2n1;
        const long long c0 = static_cast<long long>(m.y) * column_tile<kFwdTileLog2>(fs);
        float2* dst = d.work + (which ? L : 0);
        auto store = [=](long long j1, int k2, float2 v) {
            gstore(dst, static_cast<int>(j1) + (k2 << log2n1), v);
        };
        // One code path for both inputs: a is read rotated by na - 1 (src/xcorr.cu:278-282),
        // b as it is; both zero past their length. The choices are uniform per item.
        const float2* src = which ? d.b : d.a;
        const int len = which ? d.num_b : d.num_a;
        const int rot = which ? 0 : d.num_a - 1;
        column_pass_block<false, kFwdTileLog2>(
            fs, c0,
            [=](long long j1, int j2) {
                int k = static_cast<int>(j1) + (j2 << log2n1) + rot;
                k = k >= L ? k - L : k;
                return k < len ? gload(src, k) : make_float2(0.0f, 0.0f);
            },
            store, lds, table);
        __syncthreads();  // the next item reuses the LDS
    });
}

// The items a pass visits: items begin .. begin + total - 1 of the item map, or, with
// `list` set, the 2^log2_items items of each of the pairs list[0 .. *list_count - 1] (pairs of
// one shape; the count is read on the device, so the launch needs no host round trip).
struct ItemSet {
    const int* list;
    const int* list_count;
    int log2_items;
};

template <class R>
__global__ __launch_bounds__(kItemThreads) void batch_column_inv_peak_kernel(
    const PairDesc* __restrict__ pairs, const int2* __restrict__ map, int begin, int total,
    R* results, ItemSet set, const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    __shared__ unsigned long long wave_best[kItemThreads / 64];
    const int visits = set.list != nullptr ? (*set.list_count << set.log2_items) : total;
    xcd_major_items(visits, [&](int local) {
        int p;
        int within;
        if (set.list != nullptr) {
            p = set.list[local >> set.log2_items];
            within = local & ((1 << set.log2_items) - 1);
        } else {
            const int2 m = map[begin + local];
            p = m.x;
            within = m.y;
        }
        const PairDesc d = pairs[p];
        const FourStep fs = shape_of(d);
        // Indices below L <= 2^24: 32-bit arithmetic throughout.
        const int L = 1 << d.log2L;
        const int log2n1 = d.log2n1;
        const int n2 = 1 << d.log2n2;
        const int n8 = n2 >> 3;
        const int lc = column_tile_log2(fs);
        const int C = 1 << lc;
        const int slots = column_slots(n2);
        const int c0 = within * C;
        const float2* src = d.work + L;
        int tid = threadIdx.x;  // through an empty asm: no per-lane indices hoisted out of the walk
        asm volatile("" : "+v"(tid));
        // C n2 = 4096 values, 8 per lane: column c0 + tid mod C, rows t + n8 r (t = tid / C), all
        // loads in flight before the transform, which runs from registers to registers
        // (stockham_column, as in column_pass_block)
        const int col = tid & (C - 1);
        const int t = tid >> lc;
        const int j1 = c0 + col;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            v[r] = gload(src, j1 + ((t + r * n8) << log2n1));
        }
        const int lr = stockham_column<true>(lds + col * slots, v, d.log2n2, t, true, table);
        const long long lags = static_cast<long long>(d.num_a) + d.num_b - 1;
        unsigned long long best = 0ULL;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int j2 = t + stockham_order(lr, q) * n8;
            const int n = j1 + (j2 << log2n1);
            if (n < lags) {
                const float2 val = v[q];
                // |c|^2 orders the lags like |c| (ties aside), as in the 2^15 kernel; the square
                // root is taken for the peak only (batch_finish_kernel)
                const unsigned long long key = peak_key(val.x * val.x + val.y * val.y, n);
                best = key > best ? key : best;
                if (d.r != nullptr) {
                    gstore(d.r, n, val);
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long other = __shfl_xor(best, off, 64);
            best = other > best ? other : best;
        }
        if ((threadIdx.x & 63) == 0) {
            wave_best[threadIdx.x >> 6] = best;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kItemThreads / 64; ++w) {
                best = wave_best[w] > best ? wave_best[w] : best;
            }
            atomicMax(reinterpret_cast<unsigned long long*>(&results[p]), best);
        }
        __syncthreads();
    });
}

// r[n] from column j1 of the pass-2 output: sum_k2 Z[j1 + n1 k2] exp(+2 pi i j2 k2 / n2),
// one wave, float32 terms summed in float64.
__device__ double lag_magnitude(const PairDesc& d, const FourStep& fs, long long n)
{
    const long long L = 1LL << d.log2L;
    const int n1 = 1 << d.log2n1;
    const int n2 = 1 << d.log2n2;
    const long long j1 = n & (n1 - 1);
    const long long j2 = n >> d.log2n1;
    const float2* col = d.work + L + j1;
    double re = 0.0;
    double im = 0.0;
    for (int k2 = threadIdx.x; k2 < n2; k2 += 64) {
        const long long m = ((j2 * k2) & (n2 - 1)) * n1;
        const float2 w = four_step_twiddle<true>(fs, m);
        const float2 z = col[static_cast<long long>(k2) * n1];
        re += static_cast<double>(z.x) * w.x - static_cast<double>(z.y) * w.y;
        im += static_cast<double>(z.x) * w.y + static_cast<double>(z.y) * w.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        re += __shfl_xor(re, off, 64);
        im += __shfl_xor(im, off, 64);
    }
    return sqrt(re * re + im * im);
}

template <class R>
__global__ __launch_bounds__(64) void batch_finish_kernel(const PairDesc* __restrict__ pairs,
                                                          R* results,
                                                          const int* __restrict__ active)
{
    if (active != nullptr && active[blockIdx.x] == 0) {
        return;
    }
    const PairDesc d = pairs[blockIdx.x];
    if (d.fused != 0) {
        return;  // finished by its own kernel
    }
    const FourStep fs = shape_of(d);
    R* res = &results[blockIdx.x];
    const unsigned long long key = *reinterpret_cast<const unsigned long long*>(res);
    const long long lags = static_cast<long long>(d.num_a) + d.num_b - 1;
    const int idx = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(key & 0xFFFFFFFFULL));
    const float mag = sqrtf(__uint_as_float(static_cast<unsigned>(key >> 32)));  // key: |c|^2
    double frac = 0.0;
    if (idx > 0 && idx + 1 < lags) {
        const double ml = lag_magnitude(d, fs, idx - 1);
        const double mh = lag_magnitude(d, fs, idx + 1);
        const double m0 = static_cast<double>(mag);
        if (m0 > 0.0) {
            frac = mh >= ml ? mh / (m0 + mh) : -ml / (m0 + ml);
            frac = frac > 0.5 ? 0.5 : (frac < -0.5 ? -0.5 : frac);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        res->index = idx;
        res->magnitude = mag;
        res->frac = frac;
    }
}

// C * column_slots(n2) for a tile of 2^T values (n2 >= 8): 2^T + 2^T / 8 + C, C <= 2^T / 8.
template <int T>
constexpr std::size_t tile_lds_bytes()
{
    return static_cast<std::size_t>((1 << T) + (1 << T) / 8 + (1 << T) / 8) * sizeof(float2);
}

std::size_t column_fwd_lds_bytes()
{
    return tile_lds_bytes<kFwdTileLog2>();
}

std::size_t column_lds_bytes()
{
    return tile_lds_bytes<kTwiddleLog2>();
}

// Host check that a pair's column tiles fit the LDS the batched column kernels allocate.
void check_column_tile(const FourStep& fs)
{
    const std::size_t slots = column_slots(1 << fs.log2n2) * sizeof(float2);
    const std::size_t need = static_cast<std::size_t>(column_tile(fs)) * slots;
    const std::size_t need_fwd = static_cast<std::size_t>(column_tile<kFwdTileLog2>(fs)) * slots;
    if (need > column_lds_bytes() || need_fwd > column_fwd_lds_bytes()) {
        fail(RF_LIMIT, "column tile of " + std::to_string(need_fwd) + " bytes exceeds the "
                           + std::to_string(column_fwd_lds_bytes()) + "-byte LDS allocation");
    }
}

}  // namespace
}  // namespace rf
