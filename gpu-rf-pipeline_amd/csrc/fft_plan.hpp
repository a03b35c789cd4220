// The FFT plan object behind rf_fft_plan*, and the device tables a four-step transform needs.
#pragma once

#include "fft_passes.hpp"

struct rf_fft_plan {
    int nfft = 0;
    int batch = 0;
    int kind = 0;   // rf::PlanKind
    int log2n = 0;
    rf::FourStep fs{};
    float2* tables = nullptr;   // four-step lo/hi tables, or the DFT table
    float2* scratch = nullptr;  // four-step intermediate (nfft * batch values)
    std::size_t bytes = 0;
};

namespace rf {

enum PlanKind {
    kPlanTiny = 0,      // nfft in {1, 2, 4}: one lane per transform
    kPlanLds = 1,       // power of two, 8 .. 4096: one workgroup per transform (or several)
    kPlanFourStep = 2,  // power of two, 8192 .. 2^24: column pass + row pass
    kPlanDft = 3        // any other length: exact DFT against a float64-built table
};

// Two-level twiddle tables of a four-step transform of 2^log2L points, uploaded to `dev`
// (4096 + max(1, L / 4096) values). Returns the number of values written.
long long four_step_table_size(int log2L);
void upload_four_step_tables(int log2L, float2* dev);

// Enqueues the plan's transform (direction -1 forward, +1 inverse), unnormalised.
void fft_exec(const rf_fft_plan& plan, const float2* in, float2* out, int direction,
              hipStream_t stream);

}  // namespace rf
