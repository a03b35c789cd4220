// The library's constant tables: their layouts, and how each entry is computed (float64 sin/cos,
// rounded once to float32, or twice to FP16 hi/lo). Plain host C++ with no HIP types, so that
// both the build-time generator (table_gen.cpp) and the HIP sources include it.
//
// The generator writes every table into the shared library as initialised device arrays
// (twiddle.hip), so no table is allocated, uploaded or built at run time: any entry point may
// be the first one called on a device, also under stream capture.
#pragma once

#include <cmath>
#include <initializer_list>

namespace rf {

// exp(-2 pi i m / 4096), m < 4096: the single-workgroup FFTs (fft_core.hpp) read it with a
// stride of 4096 / P.
constexpr int kTwiddleLog2 = 12;
constexpr int kTwiddleSize = 1 << kTwiddleLog2;  // 4096

// Stage table of the 4096-point radix-16 transform (fft4096.hpp): [pass][k][r] =
// exp(-2 pi i r k / (16 span)), r < 16, k < span, for span 16 (offset 0) and 256 (offset 256).
constexpr int kStage4096Size = 16 * 16 + 256 * 16;  // 4352

// Factors of the 16384-point transform on 1024 lanes and of the 2^15-point split into two
// halves (fft16384.hpp), in float2 units, both in the device table and in LDS.
constexpr int kTw4 = 0;                     // span 4: 4 rows x 16
constexpr int kTw64 = kTw4 + 4 * 16;        // span 64: 64 rows x 17 (16 used)
constexpr int kTwHi = kTw64 + 64 * 17;      // 128: exp(-2 pi i h / 128)
constexpr int kTwLo = kTwHi + 128;          // 128: exp(-2 pi i l / 16384)
constexpr int kTwLds = kTwLo + 128;         // 1408 values copied to LDS
// Only in the device table: the fold factors of a 2^15-point transform split into two halves,
// exp(-2 pi i n / 32768) for n = t + P j = (per lane) * (per j).
constexpr int kFoldLane = kTwLds;           // 1024: exp(-2 pi i t / 32768)
constexpr int kFoldRow = kFoldLane + 1024;  // 32: exp(-2 pi i j / 64), the factor of n = 512 j
constexpr int kTw16384Size = kFoldRow + 32;

// The 64-channel DFT of the matrix-core channelizer (channelize_mfma.hip) as the real
// 128 x 128 matrix D[2r + d][2k + e] of z[k] = sum_r v[r] exp(+2 pi i k r / 64), stored as the
// B fragments of v_mfma_f32_32x32x16_f16: [wave w < 4][step s < 8][lane l < 64][q < 8], element
// row kk = 16 s + 8 (l / 32) + q (= 2r + d), column n = 32 w + l % 32 (= 2k + e). Hi table,
// then lo table (FP16 each).
constexpr int kDftChannels = 64;
constexpr int kDftSteps = 2 * kDftChannels / 16;   // 8 k-steps of 16
constexpr int kDftHalves = 4 * kDftSteps * 64 * 8; // per hi or lo

namespace tables {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// exp(-2 pi i (m * step mod n) / n), m < count, as (cos, sin) in float64. The index product is
// reduced exactly in integers first, so large indices keep full accuracy.
template <class Put>
void twiddles(long long count, long long n, long long step, Put put)
{
    for (long long m = 0; m < count; ++m) {
        const long long e = (m * step) % n;
        const double a = -kTwoPi * static_cast<double>(e) / static_cast<double>(n);
        put(m, std::cos(a), std::sin(a));
    }
}

// The 4096-point twiddle table.
template <class Put>
void twiddle4096(Put put)
{
    twiddles(kTwiddleSize, kTwiddleSize, 1, put);
}

template <class Put>
void stage4096(Put put)
{
    long long base = 0;
    for (int span : {16, 256}) {
        for (int k = 0; k < span; ++k) {
            // entries r * k of a transform of 16 * span points
            twiddles(16, 16LL * span, k, [&](long long r, double c, double s) {
                put(base + r, c, s);
            });
            base += 16;
        }
    }
}

// Entries not written are zero (the 17th slot of each span-64 row).
template <class Put>
void table16384(Put put)
{
    auto at = [&](long long offset) {
        return [&put, offset](long long m, double c, double s) { put(offset + m, c, s); };
    };
    for (int k = 0; k < 4; ++k) {
        twiddles(16, 64, k, at(kTw4 + 16 * k));        // exp(-2 pi i r k / 64)
    }
    for (int k = 0; k < 64; ++k) {
        twiddles(16, 1024, k, at(kTw64 + 17 * k));    // exp(-2 pi i r k / 1024)
    }
    twiddles(128, 128, 1, at(kTwHi));
    twiddles(128, 16384, 1, at(kTwLo));
    twiddles(1024, 32768, 1, at(kFoldLane));
    twiddles(32, 64, 1, at(kFoldRow));
}

// D[kk][n] of the channelizer DFT for fragment index idx (layout above), in float64:
// exp(+i theta), theta = 2 pi (k r mod M) / M; Re z gets Re v cos - Im v sin, Im z gets
// Re v sin + Im v cos.
inline double dft64_entry(int idx)
{
    const int q = idx % 8;
    const int l = (idx / 8) % 64;
    const int s = (idx / (8 * 64)) % kDftSteps;
    const int w = idx / (8 * 64 * kDftSteps);
    const int kk = 16 * s + 8 * (l / 32) + q;
    const int n = 32 * w + l % 32;
    const int r = kk >> 1;
    const int d = kk & 1;
    const int k = n >> 1;
    const int e = n & 1;
    const double theta =
        kTwoPi * static_cast<double>((k * r) % kDftChannels) / kDftChannels;
    const double c = std::cos(theta);
    const double sn = std::sin(theta);
    return e == 0 ? (d == 0 ? c : -sn) : (d == 0 ? sn : c);
}

}  // namespace tables
}  // namespace rf
