#pragma once

// Owning wrapper around an rf_fft_plan (reference: src/fft_plan.hpp:11-37, which wraps a cuFFT
// plan). Creating a plan is setup: build it before Bench::run().

#include "check.hpp"

#include <cstdio>

class FftPlan {
public:
    // `batch` transforms of `nfft` points each, stored back to back.
    FftPlan(int nfft, int batch) { RF_CHECK(rf_fft_plan_create(nfft, batch, &plan_)); }
    ~FftPlan()
    {
        if (rf_fft_plan_destroy(plan_) != RF_OK) {
            std::fprintf(stderr, "warning: rf_fft_plan_destroy failed: %s\n", rf_last_error());
        }
    }
    FftPlan(const FftPlan&) = delete;
    FftPlan& operator=(const FftPlan&) = delete;

    rf_fft_plan* get() const { return plan_; }

private:
    rf_fft_plan* plan_ = nullptr;
};
