#pragma once

// Device-pointer entry points with the reference's shape (src/device_ops.hpp:1-106), as thin
// inline wrappers over the C ABI. Each one only enqueues work: no allocation, no host
// transfers, no synchronisation. What changed, and why:
//   - cufftHandle becomes rf_fft_plan* and cudaStream_t becomes rf_stream (a hipStream_t);
//     every *_device call takes an optional trailing stream (the default stream if omitted),
//     so a caller can queue a whole pipeline on one stream and capture it as a hipGraph;
//   - taps are read from the caller's pointer on every launch: there is no constant bank, so
//     fir_load_const_taps only checks the reference's 8192-tap limit and fir_launch uses the
//     `taps` it is given for every variant;
//   - FirImpl gains FftOla (overlap-save with the library's FFT).

#include "check.hpp"
#include "device_buffer.hpp"
#include "fft_plan.hpp"
#include "types.hpp"

#include <cstddef>
#include <stdexcept>
#include <string>

// --- FIR ------------------------------------------------------------------------------------

enum class FirImpl { Naive = RF_FIR_NAIVE, Tiled = RF_FIR_TILED, TiledConst = RF_FIR_TILED_CONST,
                     Register = RF_FIR_REGISTER, FftOla = RF_FIR_FFT_OLA };

inline bool parse_fir_impl(const char* name, FirImpl& impl)
{
    int idx = 0;
    if (rf_fir_parse_impl(name, &idx) != RF_OK) {
        return false;
    }
    impl = static_cast<FirImpl>(idx);
    return true;
}

inline void fir_launch(FirImpl impl, const cf32* x, const float* taps, cf32* y, int num_samples,
                       int num_taps, rf_stream stream)
{
    RF_CHECK(rf_fir(static_cast<int>(impl), as_rf(x), taps, as_rf(y), num_samples, num_taps, 0,
                    stream));
}

inline void fir_load_const_taps(const float* /*taps*/, int num_taps, rf_stream /*stream*/)
{
    if (num_taps > RF_MAX_CONST_TAPS) {
        throw std::runtime_error("fir: " + std::to_string(num_taps) + " taps, more than the "
                                 + std::to_string(RF_MAX_CONST_TAPS)
                                 + " the constant-cache variants accept; use naive");
    }
}

// A filter prepared once for many blocks (rf_fir_plan_*): construction copies the device taps
// (and, for fft_ola, makes their spectrum) and synchronises; run() enqueues the filtering
// kernel only, with the bits of fir_launch.
class FirPlan {
public:
    FirPlan(FirImpl impl, const float* taps, int num_taps, rf_stream stream = nullptr)
    {
        RF_CHECK(rf_fir_plan_create(static_cast<int>(impl), taps, num_taps, 0, stream, &handle_));
    }
    ~FirPlan() { rf_fir_plan_destroy(handle_); }
    FirPlan(const FirPlan&) = delete;
    FirPlan& operator=(const FirPlan&) = delete;

    void run(const cf32* x, cf32* y, int num_samples, rf_stream stream = nullptr) const
    {
        RF_CHECK(rf_fir_plan_run(handle_, as_rf(x), as_rf(y), num_samples, stream));
    }

private:
    rf_fir_plan* handle_ = nullptr;
};

inline void fir_device(FirImpl impl, const cf32* x, const float* taps, cf32* y, int num_samples,
                       int num_taps, rf_stream stream = nullptr)
{
    fir_launch(impl, x, taps, y, num_samples, num_taps, stream);
}

inline void fir_naive_device(const cf32* x, const float* taps, cf32* y, int num_samples,
                             int num_taps, rf_stream stream = nullptr)
{
    fir_launch(FirImpl::Naive, x, taps, y, num_samples, num_taps, stream);
}

constexpr int kMaxFirBatch = RF_MAX_FIR_BATCH;

inline void fir_batched_device(FirImpl impl, const cf32* const* x, cf32* const* y,
                               const int* num_samples, int count, const float* taps,
                               int num_taps, rf_stream stream = nullptr)
{
    RF_CHECK(rf_fir_batched(static_cast<int>(impl), reinterpret_cast<const rf_cf32* const*>(x),
                            reinterpret_cast<rf_cf32* const*>(y), num_samples, count, taps,
                            num_taps, stream));
}

inline void fir_naive_batched_device(const cf32* const* x, cf32* const* y,
                                     const int* num_samples, int count, const float* taps,
                                     int num_taps, rf_stream stream = nullptr)
{
    fir_batched_device(FirImpl::Naive, x, y, num_samples, count, taps, num_taps, stream);
}

// --- magnitude --------------------------------------------------------------------------------

inline void magnitude_device(const cf32* x, float* mag, int num_samples,
                             rf_stream stream = nullptr)
{
    RF_CHECK(rf_magnitude(as_rf(x), mag, num_samples, stream));
}

// --- cross-correlation ------------------------------------------------------------------------

inline void xcorr_naive_device(const cf32* a, const cf32* b, cf32* r, int num_a, int num_b,
                               rf_stream stream = nullptr)
{
    RF_CHECK(rf_xcorr_direct(RF_XCORR_NAIVE, as_rf(a), as_rf(b), as_rf(r), num_a, num_b, stream));
}

inline void xcorr_tiled_device(const cf32* a, const cf32* b, cf32* r, int num_a, int num_b,
                               rf_stream stream = nullptr)
{
    RF_CHECK(rf_xcorr_direct(RF_XCORR_TILED, as_rf(a), as_rf(b), as_rf(r), num_a, num_b, stream));
}

// FFT cross-correlation for fixed lengths; construction allocates, run() enqueues.
class XcorrFft {
public:
    XcorrFft(int num_a, int num_b, bool phat = false)
    {
        RF_CHECK(rf_xcorr_fft_create(num_a, num_b, phat ? 1 : 0, &handle_));
    }
    ~XcorrFft() { rf_xcorr_fft_destroy(handle_); }
    XcorrFft(const XcorrFft&) = delete;
    XcorrFft& operator=(const XcorrFft&) = delete;

    void run(const cf32* a, const cf32* b, cf32* r, rf_stream stream = nullptr) const
    {
        RF_CHECK(rf_xcorr_fft_run(handle_, as_rf(a), as_rf(b), as_rf(r), stream));
    }
    std::size_t bytes() const
    {
        std::size_t n = 0;
        RF_CHECK(rf_xcorr_fft_bytes(handle_, &n));
        return n;
    }

private:
    rf_xcorr_fft* handle_ = nullptr;
};

// --- channelizer, spectrum ----------------------------------------------------------------------

inline int channelize_frames(int num_samples, int channels)
{
    return rf_channelize_frames(num_samples, channels);
}

// `plan`: channels points x channelize_frames() frames, or nullptr for the fused kernel.
inline void channelize_device(const rf_fft_plan* plan, const cf32* x, const float* taps, cf32* out,
                              int num_samples, int channels, int num_taps,
                              rf_stream stream = nullptr)
{
    RF_CHECK(rf_channelize(plan, as_rf(x), taps, as_rf(out), num_samples, channels, num_taps,
                           stream));
}

inline void spectrum_device(const rf_fft_plan* plan, const cf32* in, cf32* out,
                            rf_stream stream = nullptr)
{
    RF_CHECK(rf_spectrum(plan, as_rf(in), as_rf(out), stream));
}

// |spectrum_device| in one pass (rf_spectrum_magnitude; no reference counterpart): `mag` holds
// nfft * batch floats and the complex spectrum is never stored.
inline void spectrum_magnitude_device(const rf_fft_plan* plan, const cf32* in, float* mag,
                                      rf_stream stream = nullptr)
{
    RF_CHECK(rf_spectrum_magnitude(plan, as_rf(in), mag, stream));
}
