// Per-device twiddle tables, built once per process in float64 on the host and uploaded.
#pragma once

#include "rf_common.hpp"

namespace rf {

// exp(-2 pi i m / 4096) for m < 4096, float32, on the current device. Built on first use
// (a synchronous upload, outside any stream the caller owns) and kept until the process ends.
const float2* twiddle_table_4096();

// Twiddles of the 4096-point radix-16 transform's passes with span 16 and 256, laid out
// [pass][k][r] = exp(-2 pi i r k / (16 span)), r < 16, k < span (offsets 0 and 256; 4352
// values).
const float2* fft4096_stage_table();

// Factors of the 16384-point transform on 1024 lanes and of the 2^15-point split into two
// halves, laid out as the kTw* / kFold* offsets of fft16384.hpp (kTw16384Size values).
const float2* fft16384_table();

// Host-side float64 -> float32 table of exp(-2 pi i m / n) for m < count.
void fill_twiddles(float2* host, long long count, long long n, long long step);

}  // namespace rf
