#pragma once

// Host-vector operations with the reference's shape (src/ops.hpp:1-52): each takes host data,
// does its one-time setup, hands its upload / compute / download phases to Bench::run() and
// returns host data. FIR and cross-correlation keep their variant tables, selected by name
// (rfgpu --impl); the first entry is the default.

#include "bench.hpp"
#include "types.hpp"

#include <vector>

template <class Fn>
struct Impl {
    const char* name;
    const char* summary;
    Fn fn;
};

// Framed forward FFT: ceil(N / nfft) frames, last one zero-padded, unnormalised.
std::vector<cf32> spectrum_fft(const std::vector<cf32>& x, int nfft, Bench& bench);

// mag[n] = |x[n]|.
std::vector<float> magnitude_hip(const std::vector<cf32>& x, Bench& bench);

// y[n] = sum_k taps[k] x[n - k]; `block` threads per workgroup (a multiple of 32 up to 1024).
using FirFn = std::vector<cf32> (*)(const std::vector<cf32>& x, const std::vector<float>& taps,
                                    int block, Bench& bench);
constexpr int kDefaultFirBlock = RF_DEFAULT_FIR_BLOCK;
const std::vector<Impl<FirFn>>& fir_impls();

// Polyphase channelizer: ceil(N / channels) frames of `channels` channel samples.
std::vector<cf32> channelize_hip(const std::vector<cf32>& x, const std::vector<float>& taps,
                                 int channels, Bench& bench);

// r[k] = sum_n b[n + k] conj(a[n]), k = -(Na - 1) .. Nb - 1.
using XcorrFn = std::vector<cf32> (*)(const std::vector<cf32>& a, const std::vector<cf32>& b,
                                      Bench& bench);
const std::vector<Impl<XcorrFn>>& xcorr_impls();

// GCC-PHAT: same length and lag order as xcorr.
std::vector<cf32> gcc_phat_hip(const std::vector<cf32>& a, const std::vector<cf32>& b,
                               Bench& bench);

// GCC-PHAT with the peak searched on the device: the lag of the first maximum of |r|.
int tdoa_hip(const std::vector<cf32>& a, const std::vector<cf32>& b);

// The lag of every channel 1.. against channel 0 (tdoa_hip of each pair), from one batched
// GCC-PHAT over all pairs (rf_gcc_batch); pairs too short for the batch (fewer than 2049
// output lags) run tdoa_hip.
std::vector<int> tdoa_batch_hip(const std::vector<std::vector<cf32>>& channels);

// Source compatibility with callers written against the reference's names.
inline std::vector<cf32> spectrum_cufft(const std::vector<cf32>& x, int nfft, Bench& bench)
{
    return spectrum_fft(x, nfft, bench);
}
inline std::vector<float> magnitude_cuda(const std::vector<cf32>& x, Bench& bench)
{
    return magnitude_hip(x, bench);
}
inline std::vector<cf32> channelize_cuda(const std::vector<cf32>& x,
                                         const std::vector<float>& taps, int channels,
                                         Bench& bench)
{
    return channelize_hip(x, taps, channels, bench);
}
inline std::vector<cf32> gcc_phat_cuda(const std::vector<cf32>& a, const std::vector<cf32>& b,
                                       Bench& bench)
{
    return gcc_phat_hip(a, b, bench);
}
