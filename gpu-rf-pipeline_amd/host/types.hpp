#pragma once

// One complex float32 sample, I then Q: the .cf32 file layout and rf_cf32 of the C ABI
// (reference: src/types.hpp:7-12).
#include "rfhip.h"

struct cf32 {
    float re;
    float im;
};

static_assert(sizeof(cf32) == 2 * sizeof(float), "cf32 must be two packed floats");
static_assert(sizeof(cf32) == sizeof(rf_cf32), "cf32 must match rf_cf32");

inline const rf_cf32* as_rf(const cf32* p)
{
    return reinterpret_cast<const rf_cf32*>(p);
}

inline rf_cf32* as_rf(cf32* p)
{
    return reinterpret_cast<rf_cf32*>(p);
}
