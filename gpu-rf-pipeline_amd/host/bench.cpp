#include "bench.hpp"

#include "check.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

namespace {

class Event {
public:
    Event() { RF_CHECK(rf_event_create(&event_)); }
    ~Event() { rf_event_destroy(event_); }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    void record() { RF_CHECK(rf_event_record(event_, nullptr)); }
    void wait() { RF_CHECK(rf_event_synchronize(event_)); }
    double ms_since(const Event& start) const
    {
        float ms = 0.0f;
        RF_CHECK(rf_event_elapsed_ms(start.event_, event_, &ms));
        return ms;
    }

private:
    void* event_ = nullptr;
};

PhaseTime summarize(const std::vector<double>& samples)
{
    PhaseTime t;
    if (samples.empty()) {
        return t;
    }
    double sum = 0.0;
    for (double s : samples) {
        sum += s;
    }
    t.mean_ms = sum / static_cast<double>(samples.size());
    if (samples.size() > 1) {
        double sq = 0.0;
        for (double s : samples) {
            sq += (s - t.mean_ms) * (s - t.mean_ms);
        }
        t.std_ms = std::sqrt(sq / static_cast<double>(samples.size() - 1));
    }
    return t;
}

}  // namespace

void Bench::run(const std::function<void()>& upload, const std::function<void()>& compute,
                const std::function<void()>& download)
{
    for (int i = 0; i < config_.warmup; ++i) {
        upload();
        compute();
        download();
    }
    Event start, uploaded, computed, downloaded;
    std::vector<double> h2d, kernel, d2h;
    for (int i = 0; i < config_.runs; ++i) {
        start.record();
        upload();
        uploaded.record();
        compute();
        computed.record();
        download();
        downloaded.record();
        downloaded.wait();
        h2d.push_back(uploaded.ms_since(start));
        kernel.push_back(computed.ms_since(uploaded));
        d2h.push_back(downloaded.ms_since(computed));
    }
    result_.h2d = summarize(h2d);
    result_.kernel = summarize(kernel);
    result_.d2h = summarize(d2h);
}
