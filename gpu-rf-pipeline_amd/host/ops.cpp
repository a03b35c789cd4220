// Host-vector operations (ops.hpp) over the device-pointer entry points (device_ops.hpp).

#include "ops.hpp"

#include "device_ops.hpp"

#include <memory>
#include <stdexcept>

namespace {

template <int IMPL>
std::vector<cf32> fir_host(const std::vector<cf32>& x, const std::vector<float>& taps, int block,
                           Bench& bench)
{
    const int n = static_cast<int>(x.size());
    const int t = static_cast<int>(taps.size());
    std::vector<cf32> y(x.size());
    DeviceBuffer<cf32> d_x(x.size());
    DeviceBuffer<float> d_taps(taps.size());
    DeviceBuffer<cf32> d_y(y.size());
    bench.run(
        [&] {
            d_x.upload(x.data(), x.size());
            d_taps.upload(taps.data(), taps.size());
        },
        [&] {
            RF_CHECK(rf_fir(IMPL, as_rf(d_x.get()), d_taps.get(), as_rf(d_y.get()), n, t, block,
                            nullptr));
        },
        [&] { d_y.download(y.data(), y.size()); });
    return y;
}

template <int IMPL>
std::vector<cf32> xcorr_direct_host(const std::vector<cf32>& a, const std::vector<cf32>& b,
                                    Bench& bench)
{
    const int na = static_cast<int>(a.size());
    const int nb = static_cast<int>(b.size());
    std::vector<cf32> r(a.size() + b.size() - 1);
    DeviceBuffer<cf32> d_a(a.size()), d_b(b.size()), d_r(r.size());
    bench.run(
        [&] {
            d_a.upload(a.data(), a.size());
            d_b.upload(b.data(), b.size());
        },
        [&] {
            RF_CHECK(rf_xcorr_direct(IMPL, as_rf(d_a.get()), as_rf(d_b.get()), as_rf(d_r.get()),
                                     na, nb, nullptr));
        },
        [&] { d_r.download(r.data(), r.size()); });
    return r;
}

std::vector<cf32> xcorr_fft_host(const std::vector<cf32>& a, const std::vector<cf32>& b,
                                 bool phat, Bench& bench)
{
    std::vector<cf32> r(a.size() + b.size() - 1);
    DeviceBuffer<cf32> d_a(a.size()), d_b(b.size()), d_r(r.size());
    const XcorrFft h(static_cast<int>(a.size()), static_cast<int>(b.size()), phat);
    bench.run(
        [&] {
            d_a.upload(a.data(), a.size());
            d_b.upload(b.data(), b.size());
        },
        [&] { h.run(d_a.get(), d_b.get(), d_r.get()); },
        [&] { d_r.download(r.data(), r.size()); });
    return r;
}

std::vector<cf32> xcorr_fft(const std::vector<cf32>& a, const std::vector<cf32>& b, Bench& bench)
{
    return xcorr_fft_host(a, b, false, bench);
}

}  // namespace

// One host wrapper per library variant; the table's names, order and summaries come from the
// library (rf_fir_impl_name / rf_fir_impl_summary), so a new variant needs one line here.
const std::vector<Impl<FirFn>>& fir_impls()
{
    static const FirFn fns[] = {fir_host<0>, fir_host<1>, fir_host<2>, fir_host<3>,
                                fir_host<4>, fir_host<5>, fir_host<6>};
    static_assert(sizeof(fns) / sizeof(fns[0]) == RF_FIR_IMPL_COUNT, "one wrapper per variant");
    static const std::vector<Impl<FirFn>> impls = [] {
        std::vector<Impl<FirFn>> v;
        for (int i = 0; i < rf_fir_impl_count(); ++i) {
            v.push_back({rf_fir_impl_name(i), rf_fir_impl_summary(i), fns[i]});
        }
        return v;
    }();
    return impls;
}

const std::vector<Impl<XcorrFn>>& xcorr_impls()
{
    static const std::vector<Impl<XcorrFn>> impls = {
        {rf_xcorr_impl_name(RF_XCORR_NAIVE), rf_xcorr_impl_summary(RF_XCORR_NAIVE),
         xcorr_direct_host<RF_XCORR_NAIVE>},
        {rf_xcorr_impl_name(RF_XCORR_TILED), rf_xcorr_impl_summary(RF_XCORR_TILED),
         xcorr_direct_host<RF_XCORR_TILED>},
        {rf_xcorr_impl_name(RF_XCORR_FFT), rf_xcorr_impl_summary(RF_XCORR_FFT), xcorr_fft},
    };
    return impls;
}

std::vector<cf32> gcc_phat_hip(const std::vector<cf32>& a, const std::vector<cf32>& b,
                               Bench& bench)
{
    return xcorr_fft_host(a, b, true, bench);
}

int tdoa_hip(const std::vector<cf32>& a, const std::vector<cf32>& b)
{
    const std::size_t lags = a.size() + b.size() - 1;
    DeviceBuffer<cf32> d_a(a.size()), d_b(b.size()), d_r(lags);
    DeviceBuffer<rf_peak_result> d_peak(1);
    d_a.upload(a.data(), a.size());
    d_b.upload(b.data(), b.size());
    const XcorrFft h(static_cast<int>(a.size()), static_cast<int>(b.size()), true);
    h.run(d_a.get(), d_b.get(), d_r.get());
    RF_CHECK(rf_peak(as_rf(d_r.get()), static_cast<int>(lags), d_peak.get(), nullptr));
    rf_peak_result peak{};
    d_peak.download(&peak, 1);
    return peak.index - static_cast<int>(a.size() - 1);
}

std::vector<int> tdoa_batch_hip(const std::vector<std::vector<cf32>>& channels)
{
    const std::vector<cf32>& ref = channels.front();
    std::vector<int> lags(channels.size() - 1, 0);
    std::vector<std::size_t> batched;  // channel indices that go through the batch
    for (std::size_t i = 1; i < channels.size(); ++i) {
        if (ref.size() + channels[i].size() - 1 > 2048) {
            batched.push_back(i);
        } else {
            lags[i - 1] = tdoa_hip(ref, channels[i]);
        }
    }
    if (batched.empty()) {
        return lags;
    }
    const int count = static_cast<int>(batched.size());
    DeviceBuffer<cf32> d_ref(ref.size());
    d_ref.upload(ref.data(), ref.size());
    std::vector<std::unique_ptr<DeviceBuffer<cf32>>> d_ch;
    std::vector<int> na(batched.size(), static_cast<int>(ref.size()));
    std::vector<int> nb;
    std::vector<const rf_cf32*> pa(batched.size(), as_rf(d_ref.get()));
    std::vector<const rf_cf32*> pb;
    for (std::size_t i : batched) {
        d_ch.push_back(std::make_unique<DeviceBuffer<cf32>>(channels[i].size()));
        d_ch.back()->upload(channels[i].data(), channels[i].size());
        nb.push_back(static_cast<int>(channels[i].size()));
        pb.push_back(as_rf(d_ch.back()->get()));
    }
    struct BatchDeleter {
        void operator()(rf_gcc_batch* h) const { (void)rf_gcc_batch_destroy(h); }
    };
    rf_gcc_batch* raw = nullptr;
    RF_CHECK(rf_gcc_batch_create(count, na.data(), nb.data(), &raw));
    const std::unique_ptr<rf_gcc_batch, BatchDeleter> batch(raw);
    RF_CHECK(rf_gcc_batch_bind(batch.get(), pa.data(), pb.data(), nullptr));
    DeviceBuffer<rf_peak_result> d_peaks(batched.size());
    RF_CHECK(rf_gcc_batch_run(batch.get(), d_peaks.get(), nullptr));
    std::vector<rf_peak_result> peaks(batched.size());
    d_peaks.download(peaks.data(), peaks.size());
    for (std::size_t k = 0; k < batched.size(); ++k) {
        lags[batched[k] - 1] = peaks[k].index - static_cast<int>(ref.size() - 1);
    }
    return lags;
}

std::vector<cf32> spectrum_fft(const std::vector<cf32>& x, int nfft, Bench& bench)
{
    const std::size_t frame = static_cast<std::size_t>(nfft);
    const std::size_t frames = (x.size() + frame - 1) / frame;
    std::vector<cf32> out(frames * frame);
    if (frames == 0) {
        return out;
    }
    DeviceBuffer<cf32> d_in(out.size());
    DeviceBuffer<cf32> d_out(out.size());
    d_in.zero();
    const FftPlan plan(nfft, static_cast<int>(frames));
    bench.run([&] { d_in.upload(x.data(), x.size()); },
              [&] { spectrum_device(plan.get(), d_in.get(), d_out.get()); },
              [&] { d_out.download(out.data(), out.size()); });
    return out;
}

std::vector<float> magnitude_hip(const std::vector<cf32>& x, Bench& bench)
{
    std::vector<float> mag(x.size());
    DeviceBuffer<cf32> d_x(x.size());
    DeviceBuffer<float> d_m(x.size());
    bench.run([&] { d_x.upload(x.data(), x.size()); },
              [&] { magnitude_device(d_x.get(), d_m.get(), static_cast<int>(x.size())); },
              [&] { d_m.download(mag.data(), mag.size()); });
    return mag;
}

std::vector<cf32> channelize_hip(const std::vector<cf32>& x, const std::vector<float>& taps,
                                 int channels, Bench& bench)
{
    if (channels < 1) {
        throw std::runtime_error("channelize: need at least one channel");
    }
    const int n = static_cast<int>(x.size());
    const int frames = channelize_frames(n, channels);
    std::vector<cf32> z(static_cast<std::size_t>(frames) * static_cast<std::size_t>(channels));
    if (frames == 0) {
        return z;
    }
    DeviceBuffer<cf32> d_x(x.size());
    DeviceBuffer<float> d_taps(taps.size());
    DeviceBuffer<cf32> d_z(z.size());
    const FftPlan plan(channels, frames);
    bench.run(
        [&] {
            d_x.upload(x.data(), x.size());
            d_taps.upload(taps.data(), taps.size());
        },
        [&] {
            channelize_device(plan.get(), d_x.get(), d_taps.get(), d_z.get(), n, channels,
                              static_cast<int>(taps.size()));
        },
        [&] { d_z.download(z.data(), z.size()); });
    return z;
}
