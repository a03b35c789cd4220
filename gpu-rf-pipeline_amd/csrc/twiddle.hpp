// Per-device twiddle tables, built once per process in float64 on the host and uploaded.
#pragma once

#include "rf_common.hpp"

namespace rf {

// exp(-2 pi i m / 4096) for m < 4096, float32, on the current device. Built on first use
// (a synchronous upload, outside any stream the caller owns) and kept until the process ends.
const float2* twiddle_table_4096();

// Host-side float64 -> float32 table of exp(-2 pi i m / n) for m < count.
void fill_twiddles(float2* host, long long count, long long n, long long step);

}  // namespace rf
