// Pieces shared by the FIR sources (fir.hip, fir_mfma.hip). Internal to the library.
#pragma once

#include "rf_common.hpp"
#include "workspace.hpp"

namespace rf {

// Stream-ordered workspace: allocated and released on the caller's stream, so an entry point
// that needs a small table (a tap spectrum, split taps) stays enqueue-only.
struct StreamScratch {
    void* ptr = nullptr;
    hipStream_t stream = nullptr;
    StreamScratch(std::size_t bytes, hipStream_t s) : stream(s)
    {
        RF_HIP_CHECK(hipMallocAsync(&ptr, bytes, stream));
    }
    ~StreamScratch()
    {
        if (ptr != nullptr) {
            (void)hipFreeAsync(ptr, stream);
        }
    }
    StreamScratch(const StreamScratch&) = delete;
    StreamScratch& operator=(const StreamScratch&) = delete;
};

// Split-precision Toeplitz-GEMM FIR on the matrix cores (fir_mfma.hip). `workspace` holds the
// split taps: kStreamWorkspaceBytes of device memory, used in stream order.
void fir_mfma_launch(bool bf16, const float2* x, const float* taps, float2* y, int n,
                     int num_taps, hipStream_t stream, void* workspace);
// Bytes of `workspace` fir_mfma_launch uses for a filter of num_taps taps.
std::size_t fir_mfma_workspace_bytes(int num_taps);

}  // namespace rf
