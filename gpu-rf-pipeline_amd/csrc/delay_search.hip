// Learned delay search: batched GCC-PHAT over equal-length capture pairs in which a small
// network proposes the delay and the GCC-PHAT output is then evaluated exactly, but only over a
// 64-lag window around the proposal; pairs the network is unsure of take the full inverse
// transform. No single reference function: the exact path is gcc_phat_cuda
// (src/xcorr.cu:379-399) plus the peak loop of rfgpu doa (src/main.cpp:390-401).
//
// Per pair (n samples per channel, L = nextpow2(2n - 1), 2^15 <= L <= 2^20, rows of 4096):
//   1 forward column pass of a and b (shared with rf_gcc_batch);
//   2 row pass: 4096-point FFTs of a's and b's rows, PHAT weighting W[k] (the GCC-PHAT
//     cross-spectrum, 1/L included), W stored in place of b's row, and band sums of the
//     de-rotated spectrum W'[k] = W[k] exp(+2 pi i k (n-1) / L) (lag 0 at index 0) over 1024
//     bands of L/1024 bins;
//   3 features: the 1024-point inverse DFT of the band sums, c[m] = (1/B) sum over lags l = m
//     (mod 1024) of r(l) D(l), D the band's Dirichlet kernel (|D/B| >= 0.64 for |l| <= 512,
//     zero at nonzero multiples of 1024), so |c| is a coarse GCC-PHAT magnitude over lags
//     [-512, 512); feature i < 256 = max of |c| over lags -512 + 4i .. -512 + 4i + 3, divided
//     by the largest feature; the coarse proposal is the bin of the largest |c|;
//   4 proposal: the network's most probable of 64 bins of 16 lags (bin b: lags -512 + 16b ..
//     -512 + 16b + 15) and its probability, or the coarse proposal with confidence 1;
//   5 refine (confidence >= threshold): r at the 64 lags -512 + 16b - 24 .. -512 + 16b + 39
//     as a pruned inverse DFT of W, r(n) = sum_p w^(p n) sum_q W[p + R q] exp(2 pi i q n / 64),
//     w = exp(2 pi i / L), R = L / 64 (64-point transforms of W's rows seen as 64 x 64
//     matrices); the peak is the first largest |r| in the window; at either window edge the
//     pair falls back;
//   6 fallback: inverse row and column passes with the fused peak search of rf_gcc_batch.
// Peaks and the sub-sample offset follow rf_gcc_batch (first largest |r|; f = m1 / (m0 + m1)
// towards the larger neighbour).
// For L = 2^15 (n <= 2^14) steps 1-2 run as one workgroup per pair instead (gcc_pair32k.hpp:
// even and odd half-spectra of 16384 points), which stores the same W and the band sums of each
// half, and the fallback runs rf_gcc_batch's single-workgroup GCC-PHAT of those pairs.

#include "gcc_batch_kernels.hpp"
#include "gcc_pair32k.hpp"
#include "fir_internal.hpp"
#include "twiddle.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

struct rf_delay_mlp;

namespace rf {
void synth_pairs(uint64_t seed, int first, int count, int n, float2* a, float2* b, float* delays,
                 hipStream_t stream);
void delay_mlp_forward(rf_delay_mlp& h, const float* x, int batch, float* probs,
                       hipStream_t stream);
void delay_mlp_train_step(rf_delay_mlp& h, const float* x, const int* labels, int batch,
                          float lr, float* loss, hipStream_t stream);
int delay_mlp_max_batch(const rf_delay_mlp& h);
}  // namespace rf

struct rf_delay_search {
    int count = 0;
    int n = 0;
    int log2L = 0;
    rf::FourStep fs{};
    std::vector<rf::PairDesc> host;
    rf::PairDesc* dev = nullptr;
    int2* col_map = nullptr;    // col_total items: {pair, column tile} of the inverse pass
    int col_total = 0;
    int2* fwd_map = nullptr;    // fwd_total items of the forward pass (wider tiles)
    int fwd_total = 0;
    float2* work = nullptr;     // count * 2 L
    float2* tables = nullptr;   // four-step twiddles of L
    float2* partial = nullptr;  // count * n2 * 1024 band sums per row
    float2* derotate = nullptr; // L values: [row][k1] = exp(+2 pi i (row + n2 k1)(n - 1) / L)
    float* features = nullptr;  // count * 256
    int* coarse = nullptr;      // count
    float* probs = nullptr;     // count * 64
    int* active = nullptr;      // count: 1 = full inverse transform
    int* fallback = nullptr;    // count + 1: number of fallback pairs, then their indices
    bool fused = false;         // L = 2^15: one workgroup per pair (gcc_pair32k.hpp)
    float2* scratch = nullptr;  // fused: pair32k_scratch_values()
    float2* derot32k = nullptr; // fused: kDerot32kSize de-rotation factors
    std::size_t bytes = 0;
};

namespace rf {
namespace {

constexpr int kBands = 1024;
constexpr int kFeatures = RF_DELAY_FEATURES;
constexpr int kBins = RF_DELAY_BINS;
constexpr int kSpan = RF_DELAY_SPAN;
constexpr int kWindow = 64;
constexpr int kWindowMargin = 24;
constexpr int kRefineThreads = 256;
constexpr int kRefineBatch = kRefineThreads / 8;  // 64-point transforms per lds_fft call

__global__ void search_bind_kernel(PairDesc* __restrict__ pairs, int count, const float2* a,
                                   const float2* b, int n)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) {
        pairs[p].a = a + static_cast<long long>(p) * n;
        pairs[p].b = b + static_cast<long long>(p) * n;
    }
}

// Sum over each quad of lanes (4j .. 4j + 3) with DPP quad permutes (xor 1, then xor 2), which
// stay in the VALU instead of going through the LDS crossbar like a general shuffle.
__device__ __forceinline__ float quad_sum(float v)
{
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1,
                                                            0xF, 0xF, false));  // [1,0,3,2]
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E,
                                                            0xF, 0xF, false));  // [2,3,0,1]
    return v;
}

// One 4096-point row per item: W into b's row, de-rotated band sums into `partial`.
__global__ __launch_bounds__(k4096Lanes) void search_row_kernel(
    const PairDesc* __restrict__ pairs, int total, float2* __restrict__ partial,
    const float2* __restrict__ derotate, const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    xcd_major_items(total, [&](int item) {
        const PairDesc d = pairs[0];  // every pair has the same shape
        const int log2n2 = d.log2n2;
        const int p = item >> log2n2;
        const long long row = item & ((1 << log2n2) - 1);
        const PairDesc dp = pairs[p];
        const long long L = 1LL << dp.log2L;
        const float2* wa = dp.work + (row << k4096Log2);
        float2* wb = dp.work + L + (row << k4096Log2);
        const int t = threadIdx.x;
        float2 ua[16];
        float2 ub[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            ua[r] = gload(wa, t + r * k4096Lanes);
            ub[r] = gload(wb, t + r * k4096Lanes);
        }
        fft4096<false>(buf, t, ua, stages);
        __syncthreads();
        fft4096<false>(buf, t, ub, stages);
        const float scale = 1.0f / static_cast<float>(L);
        float2* out = partial + (static_cast<long long>(p) << log2n2 << 10) + (row << 10);
        const float2* drow = derotate + (row << k4096Log2);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int k1 = t + r * k4096Lanes;
            const float2 w = weigh(ua[r], ub[r], scale, 1);
            gstore(wb, k1, w);
            float2 v = cmul(w, drow[k1]);  // exp(+2 pi i k (n-1) / L), k = row + n2 k1
            // band k1 / 4: lanes 4j .. 4j + 3 hold its four bins of this row
            v.x = quad_sum(v.x);
            v.y = quad_sum(v.y);
            if ((t & 3) == 0) {
                out[k1 >> 2] = v;
            }
        }
        __syncthreads();  // the next item reuses buf
    });
}

// One pair per workgroup: band sums -> coarse correlation -> 256 features, coarse proposal.
__global__ __launch_bounds__(256) void search_features_kernel(
    const float2* __restrict__ partial, int log2n2, float* __restrict__ features,
    int* __restrict__ coarse, const float2* __restrict__ table)
{
    __shared__ float2 lds[lds_slots(kBands)];
    __shared__ unsigned long long wave_key[4];
    __shared__ float wave_max[4];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int rows = 1 << log2n2;
    const float2* src = partial + (static_cast<long long>(p) << log2n2 << 10);
    for (int j = t; j < kBands; j += 256) {
        float2 y = make_float2(0.0f, 0.0f);
        for (int r = 0; r < rows; ++r) {
            y = cadd(y, src[(r << 10) + j]);
        }
        lds[lds_pad(j)] = y;
    }
    __syncthreads();
    lds_fft<true>(lds, 10, t, t < kBands / 8, table);
    // lags -512 + 4t + s, s < 4, at index (lag mod 1024)
    float f = 0.0f;
    unsigned long long key = 0ULL;
    for (int s = 0; s < 4; ++s) {
        const int lag = -kSpan + 4 * t + s;
        const float2 c = lds[lds_pad(lag & (kBands - 1))];
        const float mag = hypotf(c.x, c.y);
        f = fmaxf(f, mag);
        const unsigned long long k = peak_key(mag, lag + kSpan);
        key = k > key ? k : key;
    }
    float fmax = f;
    for (int off = 32; off > 0; off >>= 1) {
        fmax = fmaxf(fmax, __shfl_xor(fmax, off, 64));
        const unsigned long long other = __shfl_xor(key, off, 64);
        key = other > key ? other : key;
    }
    if ((t & 63) == 0) {
        wave_max[t >> 6] = fmax;
        wave_key[t >> 6] = key;
    }
    __syncthreads();
    fmax = fmaxf(fmaxf(wave_max[0], wave_max[1]), fmaxf(wave_max[2], wave_max[3]));
    features[static_cast<long long>(p) * kFeatures + t] = fmax > 0.0f ? f / fmax : 0.0f;
    if (t == 0) {
        for (int w = 1; w < 4; ++w) {
            key = wave_key[w] > key ? wave_key[w] : key;
        }
        const int index = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(key & 0xFFFFFFFFULL));
        coarse[p] = index >> 4;  // (lag + 512) / 16
    }
}

// Sum over each row of 16 lanes (DPP rows: lanes 16 i .. 16 i + 15), left in all 16 lanes.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL,
                                                              0xF, 0xF, false));
}

__device__ __forceinline__ float row16_sum(float v)
{
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x124>(v);  // row_ror:4
    v += dpp_move<0x128>(v);  // row_ror:8
    return v;
}

// De-rotation factors exp(+2 pi i k (n - 1) / L) of bin k = 2 (t + 1024 j) + o of the L = 2^15
// search: lane part exp(+2 pi i 2 t (n - 1) / L) (t < 1024), then row parts
// exp(+2 pi i (2048 j + o) (n - 1) / L) for o = 0 (j < 16) and o = 1 (j < 16).
constexpr int kDerot32kRow = kPair32kLanes;
constexpr int kDerot32kSize = kDerot32kRow + 32;

// One half of the L = 2^15 forward search step of pair d: the PHAT weights of the even (ODD
// false) or odd bins of the rotated a (as the other lengths store them, rotated by the
// de-rotation factors' conjugates) into W, in the four-step order of the other lengths (bin k
// at (k mod 8) 4096 + k / 8; bin 2 k' + o, k' = t + 1024 j, at (2 (t mod 4) + o) 4096 + t / 4 +
// 256 j), and the de-rotated sums of the 1024 bands of 32 bins, 16 of them per half, into
// partial[half][band] (band (t + 1024 j) / 16: the 16 lanes of a DPP row, whose lane j keeps
// the sum of step j for one store per lane).
template <bool ODD>
__device__ __forceinline__ void search_pair32k_half(const PairDesc& d, int t, float2* u,
                                                    float2* v, float2* buf, const float2* tw,
                                                    float2 fold_t, float2 rot_t,
                                                    const float2* __restrict__ derot,
                                                    float2* __restrict__ partial,
                                                    const float2* __restrict__ table)
{
    constexpr int P = kPair32kLanes;
    constexpr int V = kPair32kHalf / P;
    pair32k_spectra<ODD, P, false>(d, t, u, v, buf, tw, fold_t, table);  // see Fft16384Layout
    const float scale = 1.0f / static_cast<float>(1 << kPair32kLog2L);
    const __amdgpu_buffer_rsrc_t W = __builtin_amdgcn_make_buffer_rsrc(
        d.work + (1 << kPair32kLog2L), static_cast<short>(0), -1, 0x00020000);
    const unsigned slot = static_cast<unsigned>(((2 * (t & 3) + (ODD ? 1 : 0)) << k4096Log2)
                                                + (t >> 2)) * 8u;
    float2 mine = make_float2(0.0f, 0.0f);  // lane t keeps the sum of band t / 16 + 64 (t % 16)
#pragma unroll
    for (int j = 0; j < V; ++j) {
        // a is not rotated here, so x is already the de-rotated spectrum (lag 0 at index 0)
        // whose band sums the features take; the refine reads the spectrum of the rotated a,
        // x exp(-2 pi i k (n - 1) / L)
        float2 x = phat_weight(u[j], v[j], scale);
        const float2 w = cmulconj(x, cmul(rot_t, derot[kDerot32kRow + (ODD ? 16 : 0) + j]));
        using u32x2 = unsigned __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, w), W, slot,
                                              static_cast<unsigned>(256 * j) * 8u, 0);
        x.x = row16_sum(x.x);
        x.y = row16_sum(x.y);
        mine = (t & 15) == j ? x : mine;
    }
    gstore(partial, (ODD ? kBands : 0) + (t >> 4) + 64 * (t & 15), mine);
}

// Steps 1-2 for L = 2^15, one workgroup per pair.
__global__ __launch_bounds__(kPair32kLanes) void search_pair32k_kernel(
    const PairDesc* __restrict__ pairs, int count, float2* __restrict__ partial,
    const float2* __restrict__ derot, const float2* __restrict__ table)
{
    constexpr int P = kPair32kLanes;
    constexpr int V = kPair32kHalf / P;
    __shared__ float2 buf[k16384Slots];
    __shared__ float2 tw[kTwLds];
    const int lane = threadIdx.x;
    fft16384_load_tables(tw, table, lane, P);
    const float2 fold_t = table[kFoldLane + lane];
    const float2 rot_t = derot[lane];
    float2 u[V];
    float2 v[V];
    __syncthreads();  // tw
    for (int p = blockIdx.x; p < count; p += gridDim.x) {
        // the lane index through an empty asm: no per-lane addresses hoisted out of the loop
        int t = lane;
        asm volatile("" : "+v"(t));
        const PairDesc d = pairs[p];
        float2* out = partial + static_cast<long long>(p) * 2 * kBands;
        search_pair32k_half<false>(d, t, u, v, buf, tw, fold_t, rot_t, derot, out, table);
        __syncthreads();
        search_pair32k_half<true>(d, t, u, v, buf, tw, fold_t, rot_t, derot, out, table);
        __syncthreads();  // the next pair reuses buf
    }
}

// Proposal and path per pair.
__global__ void search_decide_kernel(const float* __restrict__ probs,
                                     const int* __restrict__ coarse, int count, float threshold,
                                     rf_delay_result* __restrict__ results, int* __restrict__ active,
                                     int* __restrict__ fallback)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) {
        return;
    }
    int bin = coarse[p];
    float conf = 1.0f;
    if (probs != nullptr) {
        const float* q = probs + static_cast<long long>(p) * kBins;
        bin = 0;
        conf = q[0];
        for (int j = 1; j < kBins; ++j) {
            if (q[j] > conf) {
                conf = q[j];
                bin = j;
            }
        }
    }
    const int path = conf >= threshold ? 0 : 1;
    results[p].proposal = bin;
    results[p].confidence = conf;
    results[p].path = path;
    active[p] = path;
    if (path != 0) {
        fallback[1 + atomicAdd(&fallback[0], 1)] = p;
    }
}

// Radix-8 butterfly of the second pass of a 64-point inverse transform (span 8), as
// twiddle_and_dft computes it, with the factors w1, w2, w4 (already conjugated) in registers.
__device__ __forceinline__ void refine_radix8(float2* u, int k, float2 w1, float2 w2, float2 w4)
{
    if (k != 0) {
        u[1] = cmul(u[1], w1);
        const float2 w3 = cmul(w1, w2);
        u[2] = cmul(u[2], w2);
        u[3] = cmul(u[3], w3);
        u[4] = cmul(u[4], w4);
        u[5] = cmul(u[5], cmul(w1, w4));
        u[6] = cmul(u[6], cmul(w2, w4));
        u[7] = cmul(u[7], cmul(w3, w4));
    }
    dft8<true>(u);
}

// Exact GCC-PHAT over the 64-lag window of a refined pair; one workgroup per pair.
// With k = p + R q (p < R = L / 64, q < 64), w^(R q n) = exp(2 pi i q n / 64), so
//   r(n) = sum_k W[k] w^(k n) = sum_p w^(p n) U_p[n mod 64],  U_p[j] = sum_q W[p + R q] e^(2 pi i q j / 64).
// W sits in four-step order (bin k at (k mod n2) 4096 + k / n2) and R is a multiple of n2, so
// the residues p = n2 s + c (s < 64) of row c of W are its 4096 values viewed as a 64 x 64
// matrix [q][s]: each row is read once, contiguously, transformed as 64 columns of 64 points,
// and every lane accumulates one window lag over 16 of the columns.
// Lane t holds column s = t mod 64 at rows q = g + 4 i (g = t / 64, i < 16): the inputs of the
// two first-pass butterflies g and g + 4 of that column's radix-8 x radix-8 transform (the
// passes of lds_fft, same arithmetic), so the first pass runs in registers, one LDS exchange
// feeds the second (its factors held in registers), whose outputs go to LDS once for the
// accumulation. The next row's loads are issued before the current row's LDS work and every
// barrier of the row orders LDS only, so they stay in flight while the row is transformed.
__global__ __launch_bounds__(kRefineThreads) void search_refine_kernel(
    const PairDesc* __restrict__ pairs, rf_delay_result* __restrict__ results,
    int* __restrict__ active, int* __restrict__ fallback, const float2* __restrict__ table)
{
    constexpr int slots = column_slots(kWindow);  // odd stride between the 64 columns
    constexpr int kPerLane = k4096 / kRefineThreads;
    static_assert(kRefineThreads == 256 && kPerLane == 16, "4 lanes per column, 16 rows each");
    __shared__ float2 lds[kWindow * slots];
    __shared__ float2 part[4][kWindow];
    __shared__ float mag[kWindow];
    const int p = blockIdx.x;
    if (active[p] != 0) {
        return;
    }
    const int t = threadIdx.x;
    const PairDesc d = pairs[p];
    const FourStep fs = shape_of(d);
    const long long L = 1LL << d.log2L;
    const int log2n2 = d.log2n2;
    const int n2 = 1 << log2n2;
    const float2* W = d.work + L;
    const int bin = results[p].proposal;
    const long long lo = -kSpan + 16LL * bin - kWindowMargin;   // first lag of the window
    const long long n0 = (lo + d.num_a - 1 + L) & (L - 1);       // its circular index
    const int m = t & (kWindow - 1);
    const int g = t >> 6;                                         // columns 16 g .. 16 g + 15
    const long long n = (n0 + m) & (L - 1);
    const int j = static_cast<int>(n & (kWindow - 1));
    const float2 step = four_step_twiddle<true>(fs, (n << log2n2) & (L - 1));  // w^(n2 n)
    // second-pass factors of butterflies g and g + 4: conj(table[64 k]), conj(table[128 k]),
    // conj(table[256 k]) for k = g, g + 4
    auto tw = [&](int idx) {
        const float2 w = table[idx];
        return make_float2(w.x, -w.y);
    };
    const float2 a1 = tw(64 * g), a2 = tw(128 * g), a4 = tw(256 * g);
    const float2 b1 = tw(64 * (g + 4)), b2 = tw(128 * (g + 4)), b4 = tw(256 * (g + 4));
    float2* col = lds + m * slots;   // this lane's column s = m
    float2 acc = make_float2(0.0f, 0.0f);
    float2 v[kPerLane];
#pragma unroll
    for (int i = 0; i < kPerLane; ++i) {
        v[i] = gload(W, t + i * kRefineThreads);
    }
    for (int c = 0; c < n2; ++c) {
        // first pass in registers: butterfly g takes rows g + 8 r (v[2 r]), g + 4 rows
        // g + 4 + 8 r (v[2 r + 1])
        float2 ua[8];
        float2 ub[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            ua[r] = v[2 * r];
            ub[r] = v[2 * r + 1];
        }
        dft8<true>(ua);
        dft8<true>(ub);
        const long long p0 = (static_cast<long long>(16 * g) << log2n2) + c;
        float2 w = four_step_twiddle<true>(fs, (p0 * n) & (L - 1));  // w^(p n), p = n2 s + c
        {
            // unconditional (the last row reloads itself), so that the wait for w below can
            // leave these 16 loads in flight on every path
            const int cn = c + 1 < n2 ? c + 1 : c;
            const float2* next = W + (static_cast<long long>(cn) << k4096Log2);
#pragma unroll
            for (int i = 0; i < kPerLane; ++i) {
                v[i] = gload(next, t + i * kRefineThreads);
            }
        }
        // exchange: first-pass outputs at 8 k + r, second-pass inputs k + 8 r (k = g, g + 4)
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            col[lds_pad(8 * g + r)] = ua[r];
            col[lds_pad(8 * (g + 4) + r)] = ub[r];
        }
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            ua[r] = col[lds_pad(g + 8 * r)];
            ub[r] = col[lds_pad(g + 4 + 8 * r)];
        }
        refine_radix8(ua, g, a1, a2, a4);
        refine_radix8(ub, g + 4, b1, b2, b4);
        // the outputs k + 8 r overwrite exactly this lane's inputs
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            col[lds_pad(g + 8 * r)] = ua[r];
            col[lds_pad(g + 4 + 8 * r)] = ub[r];
        }
        lds_barrier();
#pragma unroll 4
        for (int s = 16 * g; s < 16 * g + 16; ++s) {
            acc = cadd(acc, cmul(lds[s * slots + lds_pad(j)], w));
            w = cmul(w, step);
        }
        lds_barrier();
    }
    part[g][m] = acc;
    __syncthreads();
    if (t < kWindow) {
        const float2 r = cadd(cadd(part[0][t], part[1][t]), cadd(part[2][t], part[3][t]));
        const float a = hypotf(r.x, r.y);
        mag[t] = a;
        unsigned long long key = peak_key(a, t);
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long other = __shfl_xor(key, off, 64);
            key = other > key ? other : key;
        }
        if (t == 0) {
            const int best = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned>(key & 0xFFFFFFFFULL));
            // The exact window peak judges the proposal. It is rejected (full inverse transform)
            // when it sits at the window's edge (the peak may lie outside the window), or when
            // it does not stand out from the PHAT noise floor: with unit-magnitude weights W of
            // 1/L each, a lag carrying no signal has E|r|^2 = 1/L and P(|r|^2 > t/L) = e^-t, so
            // all L lags stay below sqrt((ln L + 5) / L) with probability 1 - e^-5. A window peak
            // below that level is no evidence that the window holds the global peak.
            const float floor2 = (static_cast<float>(d.log2L) * 0.69314718f + 5.0f)
                                 / static_cast<float>(L);
            if (best == 0 || best == kWindow - 1 || mag[best] * mag[best] < floor2) {
                active[p] = 1;
                results[p].path = 1;
                fallback[1 + atomicAdd(&fallback[0], 1)] = p;
            } else {
                const double m0 = mag[best];
                const double ml = mag[best - 1];
                const double mh = mag[best + 1];
                double frac = 0.0;
                if (m0 > 0.0) {
                    frac = mh >= ml ? mh / (m0 + mh) : -ml / (m0 + ml);
                    frac = frac > 0.5 ? 0.5 : (frac < -0.5 ? -0.5 : frac);
                }
                results[p].index = static_cast<int>(lo + best + d.num_a - 1);
                results[p].magnitude = mag[best];
                results[p].frac = frac;
            }
        }
    }
}

// Inverse row pass of the fallback pairs: W row -> inverse 4096-point FFT -> conjugate
// four-step twiddle, in place (the input of the shared inverse column pass).
__global__ __launch_bounds__(k4096Lanes) void search_inverse_row_kernel(
    const PairDesc* __restrict__ pairs, const int* __restrict__ fallback,
    const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    const int log2n2 = pairs[0].log2n2;
    xcd_major_items(fallback[0] << log2n2, [&](int item) {
        const int p = fallback[1 + (item >> log2n2)];
        const long long row = item & ((1 << log2n2) - 1);
        const PairDesc d = pairs[p];
        const FourStep fs = shape_of(d);
        float2* wb = d.work + (1LL << d.log2L) + (row << k4096Log2);
        const int t = threadIdx.x;
        float2 u[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            u[r] = gload(wb, t + r * k4096Lanes);
        }
        fft4096<true>(buf, t, u, stages);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int j1 = t + r * k4096Lanes;
            gstore(wb, j1, cmul(u[r], four_step_twiddle<true>(fs, j1 * row)));
        }
        __syncthreads();
    });
}

// The search's first kernel: clears the result records (peak keys) and the fallback list.
__global__ void search_init_kernel(rf_delay_result* __restrict__ results, int count,
                                   int* __restrict__ fallback)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        *reinterpret_cast<unsigned long long*>(&results[i]) = 0ULL;
        results[i].frac = 0.0;
    }
    if (i == 0) {
        fallback[0] = 0;
    }
}

__global__ void search_labels_kernel(const float* __restrict__ delays, int count,
                                     int* __restrict__ labels)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < count) {
        const int lag = static_cast<int>(rintf(delays[p]));
        const int bin = (lag + kSpan) >> 4;
        labels[p] = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
    }
}

unsigned persistent_grid(long long items, int per_cu)
{
    const long long cap = static_cast<long long>(per_cu) * cu_count();
    return static_cast<unsigned>(ceil_div(std::min(items, cap), 8) * 8);
}

void features_launch(rf_delay_search& h, const float2* a, const float2* b, hipStream_t stream)
{
    const float2* table = twiddle_table_4096();
    search_bind_kernel<<<static_cast<unsigned>(ceil_div(h.count, 256)), 256, 0, stream>>>(
        h.dev, h.count, a, b, h.n);
    check_launch("search_bind_kernel");
    if (h.fused) {
        search_pair32k_kernel<<<static_cast<unsigned>(pair32k_grid(h.count)), kPair32kLanes, 0,
                                stream>>>(h.dev, h.count, h.partial, h.derot32k,
                                          fft16384_table());
        check_launch("search_pair32k_kernel");
        search_features_kernel<<<static_cast<unsigned>(h.count), 256, 0, stream>>>(
            h.partial, 1, h.features, h.coarse, table);  // two rows: even and odd bins
        check_launch("search_features_kernel");
        return;
    }
    allow_large_lds(reinterpret_cast<const void*>(batch_column_fwd_kernel),
                    column_fwd_lds_bytes());
    batch_column_fwd_kernel<<<persistent_grid(2LL * h.fwd_total, kFwdWorkgroupsPerCu),
                              kFwdThreads, column_fwd_lds_bytes(), stream>>>(
        h.dev, h.fwd_map, 0, h.fwd_total, table);
    check_launch("batch_column_fwd_kernel");
    const int rows = h.count << h.fs.log2n2;
    search_row_kernel<<<persistent_grid(rows, 2), k4096Lanes, 0, stream>>>(
        h.dev, rows, h.partial, h.derotate, fft4096_stage_table());
    check_launch("search_row_kernel");
    search_features_kernel<<<static_cast<unsigned>(h.count), 256, 0, stream>>>(
        h.partial, h.fs.log2n2, h.features, h.coarse, table);
    check_launch("search_features_kernel");
}

void search_launch(rf_delay_search& h, const float2* a, const float2* b, rf_delay_mlp* mlp,
                   float threshold, rf_delay_result* results, hipStream_t stream)
{
    const float2* table = twiddle_table_4096();
    search_init_kernel<<<static_cast<unsigned>(ceil_div(h.count, 256)), 256, 0, stream>>>(
        results, h.count, h.fallback);
    check_launch("search_init_kernel");
    features_launch(h, a, b, stream);
    if (mlp != nullptr) {
        const int chunk = delay_mlp_max_batch(*mlp);
        for (int first = 0; first < h.count; first += chunk) {
            const int rows = std::min(chunk, h.count - first);
            delay_mlp_forward(*mlp, h.features + static_cast<long long>(first) * kFeatures, rows,
                              h.probs + static_cast<long long>(first) * kBins, stream);
        }
    }
    search_decide_kernel<<<static_cast<unsigned>(ceil_div(h.count, 256)), 256, 0, stream>>>(
        mlp != nullptr ? h.probs : nullptr, h.coarse, h.count, threshold, results, h.active,
        h.fallback);
    check_launch("search_decide_kernel");
    search_refine_kernel<<<static_cast<unsigned>(h.count), kRefineThreads, 0, stream>>>(
        h.dev, results, h.active, h.fallback, table);
    check_launch("search_refine_kernel");
    if (h.fused) {
        // fallback pairs (their count stays on the device): the single-workgroup GCC-PHAT
        launch_pair32k(h.dev, h.fallback + 1, h.fallback, h.count, h.count, results, h.scratch,
                       stream);
        return;
    }
    // Fallback pairs only (their count stays on the device): grids sized for every pair.
    const int rows = h.count << h.fs.log2n2;
    search_inverse_row_kernel<<<persistent_grid(rows, 2), k4096Lanes, 0, stream>>>(
        h.dev, h.fallback, fft4096_stage_table());
    check_launch("search_inverse_row_kernel");
    const int col_items = h.col_total / h.count;
    batch_column_inv_peak_kernel<<<persistent_grid(h.col_total, 4), kItemThreads,
                                   column_lds_bytes(), stream>>>(
        h.dev, h.col_map, 0, h.col_total, results,
        ItemSet{h.fallback + 1, h.fallback, ilog2(col_items)}, table);
    check_launch("batch_column_inv_peak_kernel");
    batch_finish_kernel<<<static_cast<unsigned>(h.count), 64, 0, stream>>>(h.dev, results,
                                                                           h.active);
    check_launch("batch_finish_kernel");
}

void destroy_search(rf_delay_search* h)
{
    if (h == nullptr) {
        return;
    }
    (void)hipFree(h->dev);
    (void)hipFree(h->col_map);
    (void)hipFree(h->fwd_map);
    (void)hipFree(h->work);
    (void)hipFree(h->tables);
    (void)hipFree(h->partial);
    (void)hipFree(h->derotate);
    (void)hipFree(h->features);
    (void)hipFree(h->coarse);
    (void)hipFree(h->probs);
    (void)hipFree(h->active);
    (void)hipFree(h->fallback);
    (void)hipFree(h->scratch);
    (void)hipFree(h->derot32k);
    delete h;
}

rf_delay_search* create_search(int count, int n)
{
    if (count < 1 || n < 1) fail(RF_INVALID_ARG, "delay search: need count >= 1 and n >= 1");
    const long long lags = 2LL * n - 1;
    long long L = 1;
    while (L < lags) {
        L *= 2;
    }
    if (L < (1LL << 15) || L > (1LL << 20)) {
        fail(RF_LIMIT, "delay search: n = " + std::to_string(n) + " gives an FFT of "
                           + std::to_string(L) + " points; supported 2^15 .. 2^20 "
                           "(8193 <= n <= 524288)");
    }
    rf_delay_search* h = new rf_delay_search;
    try {
        h->count = count;
        h->n = n;
        h->log2L = ilog2(L);
        h->fs = make_four_step_shape(h->log2L);
        check_column_tile(h->fs);
        (void)twiddle_table_4096();
        (void)fft4096_stage_table();
        const long long tsize = four_step_table_size(h->log2L);
        RF_HIP_CHECK(hipMalloc(&h->tables, tsize * sizeof(float2)));
        upload_four_step_tables(h->log2L, h->tables);
        const long long work = 2LL * L * count;
        RF_HIP_CHECK(hipMalloc(&h->work, work * sizeof(float2)));
        const int n2 = 1 << h->fs.log2n2;
        {
            // de-rotation of the row pass, in its storage order, float64 on the host
            std::vector<float2> host(static_cast<std::size_t>(L));
            const long long rot = n - 1;
            const double two_pi = 6.283185307179586476925286766559;
            for (int row = 0; row < n2; ++row) {
                for (int k1 = 0; k1 < kTwiddleSize; ++k1) {
                    const long long k = row + static_cast<long long>(n2) * k1;
                    const long long m = (k % L) * (rot % L) % L;  // exact in 64 bits: L <= 2^20
                    const double a = two_pi * static_cast<double>(m) / static_cast<double>(L);
                    host[static_cast<std::size_t>(row) * kTwiddleSize + k1] =
                        make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
                }
            }
            RF_HIP_CHECK(hipMalloc(&h->derotate, static_cast<std::size_t>(L) * sizeof(float2)));
            RF_HIP_CHECK(hipMemcpy(h->derotate, host.data(), host.size() * sizeof(float2),
                                   hipMemcpyHostToDevice));
        }
        h->fused = h->log2L == kPair32kLog2L;
        if (h->fused) {
            (void)fft16384_table();
            const std::size_t scratch_bytes =
                static_cast<std::size_t>(pair32k_scratch_values()) * sizeof(float2);
            RF_HIP_CHECK(hipMalloc(&h->scratch, scratch_bytes));
            std::vector<float2> host(kDerot32kSize);
            const long long rot = n - 1;
            const double two_pi = 6.283185307179586476925286766559;
            auto factor = [&](long long k) {
                const long long m = (k % L) * (rot % L) % L;
                const double a = two_pi * static_cast<double>(m) / static_cast<double>(L);
                return make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
            };
            for (int t = 0; t < kPair32kLanes; ++t) {
                host[static_cast<std::size_t>(t)] = factor(2LL * t);
            }
            for (int o = 0; o < 2; ++o) {
                for (int j = 0; j < 16; ++j) {
                    host[static_cast<std::size_t>(kDerot32kRow + 16 * o + j)] =
                        factor(2048LL * j + o);
                }
            }
            RF_HIP_CHECK(hipMalloc(&h->derot32k, host.size() * sizeof(float2)));
            RF_HIP_CHECK(hipMemcpy(h->derot32k, host.data(), host.size() * sizeof(float2),
                                   hipMemcpyHostToDevice));
            h->bytes += scratch_bytes + host.size() * sizeof(float2);
        }
        const long long partial = static_cast<long long>(count) * n2 * kBands;
        RF_HIP_CHECK(hipMalloc(&h->partial, partial * sizeof(float2)));
        RF_HIP_CHECK(hipMalloc(&h->features, static_cast<std::size_t>(count) * kFeatures * sizeof(float)));
        RF_HIP_CHECK(hipMalloc(&h->coarse, static_cast<std::size_t>(count) * sizeof(int)));
        RF_HIP_CHECK(hipMalloc(&h->probs, static_cast<std::size_t>(count) * kBins * sizeof(float)));
        RF_HIP_CHECK(hipMalloc(&h->active, static_cast<std::size_t>(count) * sizeof(int)));
        RF_HIP_CHECK(hipMalloc(&h->fallback, (static_cast<std::size_t>(count) + 1) * sizeof(int)));
        h->host.resize(static_cast<std::size_t>(count));
        std::vector<int2> col_map;
        std::vector<int2> fwd_map;
        const int col_items = (1 << h->fs.log2n1) / column_tile(h->fs);
        const int fwd_items = (1 << h->fs.log2n1) / column_tile<kFwdTileLog2>(h->fs);
        for (int i = 0; i < count; ++i) {
            PairDesc& d = h->host[static_cast<std::size_t>(i)];
            d = PairDesc{};
            d.num_a = n;
            d.num_b = n;
            d.log2L = h->log2L;
            d.log2n1 = h->fs.log2n1;
            d.log2n2 = h->fs.log2n2;
            d.col_items = col_items;
            d.fwd_items = fwd_items;
            d.row_items = n2;
            d.work = h->work + 2LL * L * i;
            d.lo = h->tables;
            d.hi = h->tables + kTwiddleSize;
            d.fused = h->fused ? 1 : 0;
            for (int c = 0; c < col_items; ++c) {
                col_map.push_back(make_int2(i, c));
            }
            for (int c = 0; c < fwd_items; ++c) {
                fwd_map.push_back(make_int2(i, c));
            }
        }
        h->col_total = static_cast<int>(col_map.size());
        h->fwd_total = static_cast<int>(fwd_map.size());
        RF_HIP_CHECK(hipMalloc(&h->dev, h->host.size() * sizeof(PairDesc)));
        RF_HIP_CHECK(hipMemcpy(h->dev, h->host.data(), h->host.size() * sizeof(PairDesc),
                               hipMemcpyHostToDevice));
        RF_HIP_CHECK(hipMalloc(&h->col_map, col_map.size() * sizeof(int2)));
        RF_HIP_CHECK(hipMemcpy(h->col_map, col_map.data(), col_map.size() * sizeof(int2),
                               hipMemcpyHostToDevice));
        RF_HIP_CHECK(hipMalloc(&h->fwd_map, fwd_map.size() * sizeof(int2)));
        RF_HIP_CHECK(hipMemcpy(h->fwd_map, fwd_map.data(), fwd_map.size() * sizeof(int2),
                               hipMemcpyHostToDevice));
        h->bytes += (tsize + work + partial + L) * sizeof(float2)
                   + static_cast<std::size_t>(count) * (kFeatures + kBins + 2) * sizeof(float)
                   + h->host.size() * sizeof(PairDesc)
                   + (col_map.size() + fwd_map.size()) * sizeof(int2);
    } catch (...) {
        destroy_search(h);
        throw;
    }
    return h;
}

}  // namespace
}  // namespace rf

using rf::guarded;

extern "C" {

rf_status rf_delay_search_create(int count, int n, rf_delay_search** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "delay search: NULL output handle");
        *out = nullptr;
        *out = rf::create_search(count, n);
    });
}

rf_status rf_delay_search_destroy(rf_delay_search* h)
{
    return guarded([&] { rf::destroy_search(h); });
}

rf_status rf_delay_search_bytes(const rf_delay_search* h, size_t* bytes)
{
    return guarded([&] {
        if (h == nullptr || bytes == nullptr) rf::fail(RF_INVALID_ARG, "delay search: NULL");
        *bytes = h->bytes;
    });
}

rf_status rf_delay_search_features(rf_delay_search* h, const rf_cf32* a, const rf_cf32* b,
                                   float* features, int* coarse, rf_stream stream)
{
    return guarded([&] {
        if (h == nullptr || a == nullptr || b == nullptr) {
            rf::fail(RF_INVALID_ARG, "delay search: NULL handle or input");
        }
        hipStream_t s = rf::as_stream(stream);
        rf::features_launch(*h, reinterpret_cast<const float2*>(a),
                            reinterpret_cast<const float2*>(b), s);
        if (features != nullptr) {
            RF_HIP_CHECK(hipMemcpyAsync(features, h->features,
                                        static_cast<std::size_t>(h->count) * rf::kFeatures * sizeof(float),
                                        hipMemcpyDeviceToDevice, s));
        }
        if (coarse != nullptr) {
            RF_HIP_CHECK(hipMemcpyAsync(coarse, h->coarse, static_cast<std::size_t>(h->count) * sizeof(int),
                                        hipMemcpyDeviceToDevice, s));
        }
    });
}

rf_status rf_delay_search_run(rf_delay_search* h, const rf_cf32* a, const rf_cf32* b,
                              rf_delay_mlp* mlp, float threshold, rf_delay_result* results,
                              rf_stream stream)
{
    return guarded([&] {
        if (h == nullptr || a == nullptr || b == nullptr || results == nullptr) {
            rf::fail(RF_INVALID_ARG, "delay search: NULL handle, input or results");
        }
        rf::search_launch(*h, reinterpret_cast<const float2*>(a),
                          reinterpret_cast<const float2*>(b), mlp, threshold, results,
                          rf::as_stream(stream));
    });
}

rf_status rf_delay_mlp_train(rf_delay_mlp* mlp, unsigned long long seed, int pairs, int batch,
                             int n, float learning_rate, float* losses, rf_stream stream)
{
    return guarded([&] {
        if (mlp == nullptr) rf::fail(RF_INVALID_ARG, "delay mlp: NULL handle");
        if (pairs < 1 || batch < 1 || batch > rf::delay_mlp_max_batch(*mlp)) {
            rf::fail(RF_INVALID_ARG, "delay mlp train: need pairs >= 1 and 1 <= batch <= max_batch");
        }
        if (!(learning_rate > 0.0f)) rf::fail(RF_INVALID_ARG, "delay mlp: learning rate <= 0");
        hipStream_t s = rf::as_stream(stream);
        rf_delay_search* search = rf::create_search(batch, n);
        const int steps = static_cast<int>(rf::ceil_div(pairs, batch));
        float2* a = nullptr;
        float2* b = nullptr;
        float* delays = nullptr;
        int* labels = nullptr;
        float* loss = nullptr;
        auto release = [&] {
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(delays);
            (void)hipFree(labels);
            (void)hipFree(loss);
            rf::destroy_search(search);
        };
        try {
            const std::size_t samples = static_cast<std::size_t>(batch) * n;
            RF_HIP_CHECK(hipMalloc(&a, samples * sizeof(float2)));
            RF_HIP_CHECK(hipMalloc(&b, samples * sizeof(float2)));
            RF_HIP_CHECK(hipMalloc(&delays, static_cast<std::size_t>(batch) * sizeof(float)));
            RF_HIP_CHECK(hipMalloc(&labels, static_cast<std::size_t>(batch) * sizeof(int)));
            RF_HIP_CHECK(hipMalloc(&loss, static_cast<std::size_t>(steps) * sizeof(float)));
            for (int step = 0; step < steps; ++step) {
                const int rows = std::min(batch, pairs - step * batch);
                rf::synth_pairs(seed, step * batch, batch, n, a, b, delays, s);
                rf::features_launch(*search, a, b, s);
                rf::search_labels_kernel<<<static_cast<unsigned>(rf::ceil_div(batch, 256)), 256, 0,
                                           s>>>(delays, batch, labels);
                rf::check_launch("search_labels_kernel");
                rf::delay_mlp_train_step(*mlp, search->features, labels, rows, learning_rate,
                                         loss + step, s);
            }
            if (losses != nullptr) {
                RF_HIP_CHECK(hipMemcpyAsync(losses, loss, static_cast<std::size_t>(steps) * sizeof(float),
                                            hipMemcpyDeviceToHost, s));
            }
            RF_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            release();
            throw;
        }
        release();
    });
}

}  // extern "C"
