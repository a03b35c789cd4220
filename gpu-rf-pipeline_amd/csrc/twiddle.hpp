// Per-device constant tables (layouts and values: tables.hpp). They are device arrays
// initialised in the code object, so looking one up never allocates, uploads or synchronises,
// and works as the first call on a device and under stream capture.
#pragma once

#include "rf_common.hpp"
#include "tables.hpp"

namespace rf {

// exp(-2 pi i m / 4096) for m < 4096, float32, on the current device.
const float2* twiddle_table_4096();

// Twiddles of the 4096-point radix-16 transform's passes with span 16 and 256, laid out
// [pass][k][r] = exp(-2 pi i r k / (16 span)), r < 16, k < span (offsets 0 and 256; 4352
// values).
const float2* fft4096_stage_table();

// Factors of the 16384-point transform on 1024 lanes and of the 2^15-point split into two
// halves, laid out as the kTw* / kFold* offsets of tables.hpp (kTw16384Size values).
const float2* fft16384_table();

// B fragments of the 64-channel DFT, FP16 hi (kDftHalves values) then lo (tables.hpp).
const _Float16* dft64_table();

// Host-side float64 -> float32 table of exp(-2 pi i m / n) for m < count.
void fill_twiddles(float2* host, long long count, long long n, long long step);

}  // namespace rf
