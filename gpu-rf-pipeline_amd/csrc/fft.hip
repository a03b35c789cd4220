// FFT plans and execution (rf_fft_plan_*, rf_fft_exec, rf_spectrum). Replaces the reference's
// cuFFT calls (src/fft_plan.hpp:11-37, src/spectrum.cu:11-17) with the library's own
// transforms; see fft_core.hpp and fft_passes.hpp for the method.

#include "fft_plan.hpp"
#include "twiddle.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace rf {
namespace {

// Where a transform's result goes: X itself, or with MAG, |X| (cabs_f, the magnitude kernel's
// formula, so spectrum-then-magnitude and the fused store agree bit for bit) into `mag`.
template <bool MAG>
struct SpectrumOut {
    float2* out;
    float* mag;
    __device__ __forceinline__ void put(long long i, float2 v) const
    {
        if constexpr (MAG) {
            mag[i] = cabs_f(v);
        } else {
            out[i] = v;
        }
    }
};

// nfft in {1, 2, 4}: one lane per transform, in registers.
template <bool INV, bool MAG>
__global__ void fft_tiny_kernel(const float2* in, SpectrumOut<MAG> dst, int n, long long batch)
{
    const long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= batch) {
        return;
    }
    const float2* src = in + f * n;
    const long long o = f * n;
    if (n == 1) {
        dst.put(o, src[0]);
    } else if (n == 2) {
        float2 a = src[0];
        float2 b = src[1];
        dft2<INV>(a, b);
        dst.put(o, a);
        dst.put(o + 1, b);
    } else {
        float2 v0 = src[0], v1 = src[1], v2 = src[2], v3 = src[3];
        dft4<INV>(v0, v1, v2, v3);
        dst.put(o, v0);
        dst.put(o + 1, v1);
        dst.put(o + 2, v2);
        dst.put(o + 3, v3);
    }
}

// Exact DFT for lengths that are not powers of two: one lane per output bin, the frame read
// in order, twiddle index (j * k) mod n kept in integers, table built in float64.
template <bool INV, bool MAG>
__global__ void dft_kernel(const float2* __restrict__ in, SpectrumOut<MAG> out, int n,
                           long long batch, const float2* __restrict__ table)
{
    const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= batch * n) {
        return;
    }
    const long long f = idx / n;
    const int k = static_cast<int>(idx - f * n);
    const float2* src = in + f * n;
    float re = 0.0f;
    float im = 0.0f;
    int e = 0;
    for (int j = 0; j < n; ++j) {
        float2 w = table[e];
        if (INV) {
            w.y = -w.y;
        }
        const float2 v = src[j];
        re = fmaf(v.x, w.x, re);
        re = fmaf(-v.y, w.y, re);
        im = fmaf(v.x, w.y, im);
        im = fmaf(v.y, w.x, im);
        e += k;
        if (e >= n) {
            e -= n;
        }
    }
    out.put(idx, make_float2(re, im));
}

// Row pass of the four-step transform for one transform of the batch, natural-order output:
// row k2 (n1 contiguous points) -> X[k2 + n2 * k1].
template <bool INV, bool MAG>
__global__ __launch_bounds__(512) void four_step_row_kernel(const float2* __restrict__ in,
                                                            SpectrumOut<MAG> out, FourStep fs,
                                                            const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const int n1 = 1 << fs.log2n1;
    const int n2 = 1 << fs.log2n2;
    const int F = lds_batch_per_block(fs.log2n1);
    const int slots = lds_slots(n1);
    const long long L = 1LL << fs.log2L;
    const long long transform = blockIdx.y;
    const int row0 = blockIdx.x * F;
    const int count = n2 - row0 < F ? n2 - row0 : F;
    const float2* src = in + transform * L + static_cast<long long>(row0) * n1;
    const int total = count * n1;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        lds[(e >> fs.log2n1) * slots + lds_pad(e & (n1 - 1))] = src[e];
    }
    __syncthreads();
    const int per = n1 >> 3;
    const int f = threadIdx.x / per;
    const int t = threadIdx.x - f * per;
    lds_fft<INV>(lds + (f < count ? f : 0) * slots, fs.log2n1, t, f < count, table);
    const long long o = transform * L;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int r = e >> fs.log2n1;
        const int k1 = e & (n1 - 1);
        out.put(o + (row0 + r) + static_cast<long long>(n2) * k1, lds[r * slots + lds_pad(k1)]);
    }
}

template <bool INV>
__global__ __launch_bounds__(512) void four_step_column_kernel(const float2* __restrict__ in,
                                                               float2* __restrict__ out,
                                                               FourStep fs,
                                                               const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const long long L = 1LL << fs.log2L;
    const long long transform = blockIdx.y;
    const float2* src = in + transform * L;
    float2* dst = out + transform * L;
    const int log2n1 = fs.log2n1;
    const long long c0 = static_cast<long long>(blockIdx.x) * column_tile(fs);
    column_pass_block<INV>(
        fs, c0,
        [=](long long j1, int j2) { return src[j1 + (static_cast<long long>(j2) << log2n1)]; },
        [=](long long j1, int k2, float2 v) { dst[j1 + (static_cast<long long>(k2) << log2n1)] = v; },
        lds, table);
}

// The same batched transform with |X| (cabs_f) as the store epilogue (the magnitude kernel's
// formula, so the result equals spectrum-then-magnitude bit for bit).
__global__ __launch_bounds__(512) void fft_lds_batch_magnitude_kernel(
    const float2* in, float* out, int log2P, long long batch, const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    const int P = 1 << log2P;
    const int F = lds_batch_per_block(log2P);
    const int slots = lds_slots(P);
    const long long first = static_cast<long long>(blockIdx.x) * F;
    const int count = static_cast<int>(batch - first < F ? batch - first : F);
    const long long base = first * P;
    const int total = count * P;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int f = e >> log2P;
        lds[f * slots + lds_pad(e & (P - 1))] = in[base + e];
    }
    __syncthreads();
    const int per = P >> 3;
    const int f = threadIdx.x / per;
    const int t = threadIdx.x - f * per;
    lds_fft<false>(lds + (f < count ? f : 0) * slots, log2P, t, f < count, table);
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int fe = e >> log2P;
        const float2 v = lds[fe * slots + lds_pad(e & (P - 1))];
        out[base + e] = cabs_f(v);
    }
}

std::size_t lds_batch_bytes(int log2P)
{
    return static_cast<std::size_t>(lds_batch_per_block(log2P)) * lds_slots(1 << log2P)
           * sizeof(float2);
}

// The plan's transform into `out`, or with MAG its magnitude into `mag` (out unused); MAG
// only for the plan kinds other than kPlanLds, which has its own fused kernel.
template <bool INV, bool MAG = false>
void exec_dir(const rf_fft_plan& plan, const float2* in, float2* out, hipStream_t stream,
              float* mag = nullptr)
{
    const long long batch = plan.batch;
    const SpectrumOut<MAG> dst{out, mag};
    switch (plan.kind) {
    case kPlanTiny: {
        const unsigned grid = static_cast<unsigned>(ceil_div(batch, 256));
        fft_tiny_kernel<INV, MAG><<<grid, 256, 0, stream>>>(in, dst, plan.nfft, batch);
        check_launch("fft_tiny_kernel");
        return;
    }
    case kPlanLds: {
        if constexpr (MAG) {
            fail(RF_INVALID_ARG, "exec_dir: LDS plans have their own magnitude kernel");
        }
        const int F = lds_batch_per_block(plan.log2n);
        const int threads = F * (plan.nfft >> 3);
        const unsigned grid = static_cast<unsigned>(ceil_div(batch, F));
        fft_lds_batch_kernel<INV><<<grid, threads, lds_batch_bytes(plan.log2n), stream>>>(
            in, out, plan.log2n, batch, 1.0f, twiddle_table_4096());
        check_launch("fft_lds_batch_kernel");
        return;
    }
    case kPlanFourStep: {
        const FourStep& fs = plan.fs;
        const float2* table = twiddle_table_4096();
        const int C = column_tile(fs);
        const dim3 cgrid(static_cast<unsigned>((1 << fs.log2n1) / C), static_cast<unsigned>(batch));
        const std::size_t cbytes =
            static_cast<std::size_t>(C) * column_slots(1 << fs.log2n2) * sizeof(float2);
        four_step_column_kernel<INV><<<cgrid, C * (1 << fs.log2n2) / 8, cbytes, stream>>>(
            in, plan.scratch, fs, table);
        check_launch("four_step_column_kernel");
        const int F = lds_batch_per_block(fs.log2n1);
        const dim3 rgrid(static_cast<unsigned>(ceil_div(1 << fs.log2n2, F)),
                         static_cast<unsigned>(batch));
        four_step_row_kernel<INV, MAG><<<rgrid, F * ((1 << fs.log2n1) >> 3),
                                         lds_batch_bytes(fs.log2n1), stream>>>(plan.scratch, dst,
                                                                               fs, table);
        check_launch("four_step_row_kernel");
        return;
    }
    case kPlanDft: {
        const long long total = batch * plan.nfft;
        const unsigned grid = static_cast<unsigned>(ceil_div(total, 256));
        if (!MAG && in == out) {
            // The DFT reads whole frames while writing bins, so it needs a separate output.
            dft_kernel<INV, false><<<grid, 256, 0, stream>>>(
                in, SpectrumOut<false>{plan.scratch, nullptr}, plan.nfft, batch, plan.tables);
            check_launch("dft_kernel");
            RF_HIP_CHECK(hipMemcpyAsync(out, plan.scratch, total * sizeof(float2),
                                        hipMemcpyDeviceToDevice, stream));
        } else {
            dft_kernel<INV, MAG><<<grid, 256, 0, stream>>>(in, dst, plan.nfft, batch,
                                                           plan.tables);
            check_launch("dft_kernel");
        }
        return;
    }
    }
}

}  // namespace

long long four_step_table_size(int log2L)
{
    const long long hi = log2L > kTwiddleLog2 ? (1LL << (log2L - kTwiddleLog2)) : 1;
    return kTwiddleSize + hi;
}

void upload_four_step_tables(int log2L, float2* dev)
{
    const long long L = 1LL << log2L;
    const long long hi = four_step_table_size(log2L) - kTwiddleSize;
    std::vector<float2> host(static_cast<std::size_t>(kTwiddleSize + hi));
    fill_twiddles(host.data(), kTwiddleSize, L, 1);
    fill_twiddles(host.data() + kTwiddleSize, hi, L, kTwiddleSize);
    RF_HIP_CHECK(hipMemcpy(dev, host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice));
}

void fft_exec(const rf_fft_plan& plan, const float2* in, float2* out, int direction,
              hipStream_t stream)
{
    if (plan.batch == 0) {
        return;
    }
    if (direction == RF_FFT_INVERSE) {
        exec_dir<true>(plan, in, out, stream);
    } else {
        exec_dir<false>(plan, in, out, stream);
    }
}

}  // namespace rf

using rf::guarded;

namespace {

void destroy_plan(rf_fft_plan* plan)
{
    if (plan != nullptr) {
        (void)hipFree(plan->tables);
        (void)hipFree(plan->scratch);
        delete plan;
    }
}

}  // namespace

extern "C" {

rf_status rf_fft_plan_create(int nfft, int batch, rf_fft_plan** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "fft plan: NULL output handle");
        *out = nullptr;
        if (nfft < 1 || batch < 0) {
            rf::fail(RF_INVALID_ARG, "fft plan: need nfft >= 1 and batch >= 0");
        }
        if (rf::is_pow2(nfft) && nfft > (1 << 24)) {
            rf::fail(RF_LIMIT, "fft plan: " + std::to_string(nfft)
                                   + " points, more than the supported 2^24");
        }
        rf_fft_plan* plan = new rf_fft_plan;
        plan->nfft = nfft;
        plan->batch = batch;
        try {
            (void)rf::twiddle_table_4096();
            const long long total = static_cast<long long>(nfft) * batch;
            if (nfft <= 4 && rf::is_pow2(nfft)) {
                plan->kind = rf::kPlanTiny;
            } else if (rf::is_pow2(nfft) && nfft <= rf::kMaxLdsFft) {
                plan->kind = rf::kPlanLds;
                plan->log2n = rf::ilog2(nfft);
            } else if (rf::is_pow2(nfft)) {
                plan->log2n = rf::ilog2(nfft);
                plan->kind = rf::kPlanFourStep;
                plan->fs = rf::make_four_step_shape(plan->log2n);
                const long long tsize = rf::four_step_table_size(plan->log2n);
                RF_HIP_CHECK(hipMalloc(&plan->tables, tsize * sizeof(float2)));
                rf::upload_four_step_tables(plan->log2n, plan->tables);
                plan->fs.lo = plan->tables;
                plan->fs.hi = plan->tables + rf::kTwiddleSize;
                if (total > 0) {
                    RF_HIP_CHECK(hipMalloc(&plan->scratch, total * sizeof(float2)));
                }
                plan->bytes = (tsize + total) * sizeof(float2);
            } else {
                plan->kind = rf::kPlanDft;
                std::vector<float2> host(static_cast<std::size_t>(nfft));
                rf::fill_twiddles(host.data(), nfft, nfft, 1);
                RF_HIP_CHECK(hipMalloc(&plan->tables, host.size() * sizeof(float2)));
                RF_HIP_CHECK(hipMemcpy(plan->tables, host.data(), host.size() * sizeof(float2),
                                       hipMemcpyHostToDevice));
                if (total > 0) {
                    RF_HIP_CHECK(hipMalloc(&plan->scratch, total * sizeof(float2)));
                }
                plan->bytes = (static_cast<std::size_t>(nfft) + total) * sizeof(float2);
            }
        } catch (...) {
            destroy_plan(plan);
            throw;
        }
        *out = plan;
    });
}

rf_status rf_fft_plan_destroy(rf_fft_plan* plan)
{
    return guarded([&] { destroy_plan(plan); });
}

rf_status rf_fft_plan_bytes(const rf_fft_plan* plan, size_t* bytes)
{
    return guarded([&] {
        if (plan == nullptr || bytes == nullptr) rf::fail(RF_INVALID_ARG, "fft plan: NULL");
        *bytes = plan->bytes;
    });
}

rf_status rf_fft_exec(const rf_fft_plan* plan, const rf_cf32* in, rf_cf32* out, int direction,
                      rf_stream stream)
{
    return guarded([&] {
        if (plan == nullptr) rf::fail(RF_INVALID_ARG, "fft exec: NULL plan");
        if (direction != RF_FFT_FORWARD && direction != RF_FFT_INVERSE) {
            rf::fail(RF_INVALID_ARG, "fft exec: direction must be -1 or +1");
        }
        if (plan->batch > 0 && (in == nullptr || out == nullptr)) {
            rf::fail(RF_INVALID_ARG, "fft exec: NULL buffer");
        }
        rf::fft_exec(*plan, reinterpret_cast<const float2*>(in), reinterpret_cast<float2*>(out),
                     direction, rf::as_stream(stream));
    });
}

rf_status rf_spectrum(const rf_fft_plan* plan, const rf_cf32* in, rf_cf32* out,
                      rf_stream stream)
{
    return rf_fft_exec(plan, in, out, RF_FFT_FORWARD, stream);
}

rf_status rf_spectrum_magnitude(const rf_fft_plan* plan, const rf_cf32* in, float* mag,
                                rf_stream stream)
{
    return guarded([&] {
        if (plan == nullptr || in == nullptr || mag == nullptr) {
            rf::fail(RF_INVALID_ARG, "spectrum magnitude: NULL plan or buffer");
        }
        if (plan->batch == 0) {
            return;
        }
        hipStream_t s = rf::as_stream(stream);
        if (plan->kind == rf::kPlanLds) {
            const int F = rf::lds_batch_per_block(plan->log2n);
            const int threads = F * (plan->nfft >> 3);
            const unsigned grid = static_cast<unsigned>(rf::ceil_div(plan->batch, F));
            rf::fft_lds_batch_magnitude_kernel<<<grid, threads, rf::lds_batch_bytes(plan->log2n),
                                                 s>>>(reinterpret_cast<const float2*>(in), mag,
                                                      plan->log2n, plan->batch,
                                                      rf::twiddle_table_4096());
            rf::check_launch("fft_lds_batch_magnitude_kernel");
            return;
        }
        // Other lengths: the plan's last pass (row pass, DFT, tiny transform) stores |X|.
        rf::exec_dir<false, true>(*plan, reinterpret_cast<const float2*>(in), nullptr, s, mag);
    });
}

}  // extern "C"
