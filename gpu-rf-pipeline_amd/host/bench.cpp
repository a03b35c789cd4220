// Bench::run (bench.hpp). Every timed run gets its own four device events, created before the
// first run; the loop only records them, and one synchronisation on the last event ends the
// measurement, so the host never stalls between runs for timing. Phase times are then read
// from consecutive event pairs and summarised with Welford's running mean and variance.

#include "bench.hpp"

#include "check.hpp"

#include <cmath>
#include <vector>

namespace {

// A fixed set of device events, destroyed with the object.
class EventPool {
public:
    EventPool(std::size_t count, rf_stream stream) : events_(count, nullptr), stream_(stream)
    {
        for (void*& e : events_) {
            RF_CHECK(rf_event_create(&e));
        }
    }
    ~EventPool()
    {
        for (void* e : events_) {
            if (e != nullptr) {
                (void)rf_event_destroy(e);
            }
        }
    }
    EventPool(const EventPool&) = delete;
    EventPool& operator=(const EventPool&) = delete;

    void mark(std::size_t i) { RF_CHECK(rf_event_record(events_[i], stream_)); }
    void wait(std::size_t i) { RF_CHECK(rf_event_synchronize(events_[i])); }
    double between(std::size_t from, std::size_t to) const
    {
        float ms = 0.0f;
        RF_CHECK(rf_event_elapsed_ms(events_[from], events_[to], &ms));
        return ms;
    }

private:
    std::vector<void*> events_;
    rf_stream stream_;
};

// Welford's online mean and sum of squared deviations.
class Running {
public:
    void add(double v)
    {
        ++n_;
        const double d = v - mean_;
        mean_ += d / static_cast<double>(n_);
        m2_ += d * (v - mean_);
    }
    PhaseTime time() const
    {
        PhaseTime t;
        t.mean_ms = mean_;
        t.std_ms = n_ > 1 ? std::sqrt(m2_ / static_cast<double>(n_ - 1)) : 0.0;
        return t;
    }

private:
    long n_ = 0;
    double mean_ = 0.0;
    double m2_ = 0.0;
};

}  // namespace

void Bench::run(const std::function<void()>& upload, const std::function<void()>& compute,
                const std::function<void()>& download)
{
    for (int i = 0; i < config_.warmup; ++i) {
        upload();
        compute();
        download();
    }
    const std::size_t runs = static_cast<std::size_t>(config_.runs);
    // marks of run r: 4r before upload, 4r+1 after it, 4r+2 after compute, 4r+3 after download
    EventPool marks(4 * runs, config_.stream);
    for (std::size_t r = 0; r < runs; ++r) {
        marks.mark(4 * r);
        upload();
        marks.mark(4 * r + 1);
        compute();
        marks.mark(4 * r + 2);
        download();
        marks.mark(4 * r + 3);
    }
    if (runs > 0) {
        marks.wait(4 * runs - 1);
    }
    Running h2d, kernel, d2h;
    for (std::size_t r = 0; r < runs; ++r) {
        h2d.add(marks.between(4 * r, 4 * r + 1));
        kernel.add(marks.between(4 * r + 1, 4 * r + 2));
        d2h.add(marks.between(4 * r + 2, 4 * r + 3));
    }
    result_.h2d = h2d.time();
    result_.kernel = kernel.time();
    result_.d2h = d2h.time();
}
