// Magnitude (src/magnitude.cu:17-39, tools/reference.py:98-100): mag[n] = |x[n]|.
// HBM-bound streaming kernel: each lane moves two samples per step with one 16-byte load and
// one 8-byte store; a grid of a few workgroups per CU strides over the buffer.

#include "rf_common.hpp"

namespace rf {
namespace {

__global__ __launch_bounds__(256) void magnitude_kernel(const float2* __restrict__ x,
                                                        float* __restrict__ mag, int n)
{
    const long long pairs = n / 2;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    const bool aligned = (reinterpret_cast<uintptr_t>(x) & 15) == 0
                         && (reinterpret_cast<uintptr_t>(mag) & 7) == 0;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < pairs;
         i += stride) {
        if (aligned) {
            const float4 v = reinterpret_cast<const float4*>(x)[i];
            reinterpret_cast<float2*>(mag)[i] =
                make_float2(cabs_f(make_float2(v.x, v.y)), cabs_f(make_float2(v.z, v.w)));
        } else {
            mag[2 * i] = cabs_f(x[2 * i]);
            mag[2 * i + 1] = cabs_f(x[2 * i + 1]);
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        mag[n - 1] = cabs_f(x[n - 1]);
    }
}

}  // namespace
}  // namespace rf

extern "C" rf_status rf_magnitude(const rf_cf32* x, float* mag, int num_samples,
                                  rf_stream stream)
{
    return rf::guarded([&] {
        if (num_samples < 0) rf::fail(RF_INVALID_ARG, "magnitude: negative sample count");
        if (num_samples == 0) return;
        if (x == nullptr || mag == nullptr) rf::fail(RF_INVALID_ARG, "magnitude: NULL buffer");
        const long long pairs = num_samples / 2 > 0 ? num_samples / 2 : 1;
        const long long cap = 8LL * rf::cu_count();
        long long grid = rf::ceil_div(pairs, 256);
        grid = grid < cap ? grid : cap;
        rf::magnitude_kernel<<<static_cast<unsigned>(grid), 256, 0, rf::as_stream(stream)>>>(
            reinterpret_cast<const float2*>(x), mag, num_samples);
        rf::check_launch("magnitude_kernel");
    });
}
