// Streaming FIR (stream.hpp) as a three-stage pipeline on HIP streams.
//
// Stages: uploads run on one stream, filtering on a second, downloads on a third, so both copy
// directions and the compute units work on different chunks at the same time. Chunk c uses slot
// c % depth (device window and output buffers, pinned staging, three events); the compute
// stream waits for the slot's upload event, the download stream for its compute event, and the
// host reuses a slot only once its previous download has completed (retire), which also orders
// the next upload after the filter that read the slot.
//
// History: each output needs the T - 1 samples before it. They are uploaded once, with their
// own chunk, and then kept on the device: after filtering a chunk, the compute stream copies the
// last min(T - 1, samples so far) samples of its window into `history`, and before filtering
// the next chunk it copies them in front of that chunk's samples. Chunk c's window is
// [history | chunk]; the filter over the window gives the chunk's outputs at offset `hist`, so
// direct variants produce exactly the bits of one call over the whole capture.
//
// The filter is prepared once (rf_fir_plan: device taps, fft_ola's spectrum), before the clock.
//
// Copy mode's host copies (pageable capture -> pinned staging, pinned staging -> pageable
// result) run on a small pool of host threads: one thread copies about 35 GB/s, a PCIe 5 DMA
// engine moves 50+ GB/s per direction, so a single-threaded copy would set the pace. Staging is
// touched once at setup: page-locked memory is not faulted in on the CPU side until first use,
// and those faults would otherwise land in the first chunks.

#include "stream.hpp"

#include "check.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

// Host threads that split one memcpy between them; copy() returns when every part is done.
class CopyPool {
public:
    explicit CopyPool(int threads)
    {
        for (int t = 0; t < threads; ++t) {
            workers_.emplace_back([this, t] { work(t); });
        }
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            stop_ = true;
        }
        wake_.notify_all();
        for (std::thread& w : workers_) {
            w.join();
        }
    }
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;

    void copy(void* dst, const void* src, std::size_t bytes)
    {
        constexpr std::size_t kMinPart = std::size_t(1) << 20;
        const std::size_t parts = std::min<std::size_t>(workers_.size() + 1,
                                                        std::max<std::size_t>(1, bytes / kMinPart));
        if (parts <= 1) {
            std::memcpy(dst, src, bytes);
            return;
        }
        const auto part = [=](std::size_t k) {
            const std::size_t b = bytes * k / parts;
            const std::size_t e = bytes * (k + 1) / parts;
            std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b);
        };
        {
            std::lock_guard<std::mutex> lock(mutex_);
            job_ = part;
            parts_ = parts;
            pending_ = parts - 1;  // parts 1.. go to the workers, part 0 stays here
            ++generation_;
        }
        wake_.notify_all();
        part(0);
        std::unique_lock<std::mutex> lock(mutex_);
        done_.wait(lock, [this] { return pending_ == 0; });
    }

private:
    void work(int t)
    {
        std::size_t seen = 0;
        for (;;) {
            std::function<void(std::size_t)> job;
            std::size_t parts = 0;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                wake_.wait(lock, [&] { return stop_ || generation_ != seen; });
                if (stop_) {
                    return;
                }
                seen = generation_;
                job = job_;
                parts = parts_;
            }
            const std::size_t k = static_cast<std::size_t>(t) + 1;
            if (k < parts) {
                job(k);
                std::lock_guard<std::mutex> lock(mutex_);
                if (--pending_ == 0) {
                    done_.notify_one();
                }
            }
        }
    }

    std::vector<std::thread> workers_;
    std::mutex mutex_;
    std::condition_variable wake_, done_;
    std::function<void(std::size_t)> job_;
    std::size_t parts_ = 0;
    std::size_t pending_ = 0;
    std::size_t generation_ = 0;
    bool stop_ = false;
};

int copy_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::min(7u, hw > 1 ? hw - 1 : 0u));
}

// --- owning handles over the C ABI ----------------------------------------------------------

struct StreamHandle {
    rf_stream s = nullptr;
    StreamHandle() { RF_CHECK(rf_stream_create(&s)); }
    ~StreamHandle() { (void)rf_stream_destroy(s); }
    StreamHandle(const StreamHandle&) = delete;
    StreamHandle& operator=(const StreamHandle&) = delete;
};

struct EventHandle {
    void* e = nullptr;
    EventHandle() { RF_CHECK(rf_event_create(&e)); }
    ~EventHandle() { (void)rf_event_destroy(e); }
    EventHandle(const EventHandle&) = delete;
    EventHandle& operator=(const EventHandle&) = delete;
};

// Device memory, or page-locked host memory (optionally mapped for the device).
class Memory {
public:
    enum class Kind { Device, Pinned, Mapped };
    Memory(Kind kind, std::size_t samples) : kind_(kind)
    {
        const std::size_t bytes = std::max<std::size_t>(samples, 1) * sizeof(cf32);
        void* p = nullptr;
        switch (kind_) {
        case Kind::Device:
            RF_CHECK(rf_malloc(&p, bytes));
            host_ = dev_ = static_cast<cf32*>(p);
            break;
        case Kind::Pinned:
            RF_CHECK(rf_host_alloc(&p, bytes));
            host_ = dev_ = static_cast<cf32*>(p);
            break;
        case Kind::Mapped: {
            void* d = nullptr;
            RF_CHECK(rf_host_alloc_mapped(&p, &d, bytes));
            host_ = static_cast<cf32*>(p);
            dev_ = static_cast<cf32*>(d);
            break;
        }
        }
    }
    ~Memory()
    {
        if (kind_ == Kind::Device) {
            (void)rf_free(dev_);
        } else {
            (void)rf_host_free(host_);
        }
    }
    Memory(const Memory&) = delete;
    Memory& operator=(const Memory&) = delete;

    cf32* host() const { return host_; }    // CPU address (pinned kinds)
    cf32* device() const { return dev_; }   // address for kernels and copies

private:
    Kind kind_;
    cf32* host_ = nullptr;
    cf32* dev_ = nullptr;
};

using Kind = Memory::Kind;

struct PlanDeleter {
    void operator()(rf_fir_plan* p) const { (void)rf_fir_plan_destroy(p); }
};
using PlanHandle = std::unique_ptr<rf_fir_plan, PlanDeleter>;

// One ring slot and the chunk it holds.
struct Slot {
    std::unique_ptr<Memory> window;    // device [history | chunk] (copy, registered)
    std::unique_ptr<Memory> out;       // filter output over the window: device, or mapped
    std::unique_ptr<Memory> stage_in;  // pinned staging (copy)
    std::unique_ptr<Memory> stage_out;
    EventHandle uploaded, filtered, downloaded;
    bool busy = false;
    std::size_t start = 0;  // first capture index of the chunk in flight
    std::size_t count = 0;  // its samples
    std::size_t hist = 0;   // history samples in front of it in the window
};

class Pipeline {
public:
    Pipeline(const std::vector<cf32>& x, const std::vector<float>& taps,
             const StreamConfig& config, std::vector<cf32>& y)
        : x_(x), y_(y), config_(config), halo_(taps.size() - 1)
    {
        const std::size_t n = x.size();
        const std::size_t chunk = static_cast<std::size_t>(config.chunk);
        const std::size_t window = chunk + halo_;
        taps_ = std::make_unique<Memory>(Kind::Device, taps.size());
        RF_CHECK(rf_memcpy_h2d(taps_->device(), taps.data(), taps.size() * sizeof(float),
                               nullptr));
        RF_CHECK(rf_stream_synchronize(nullptr));
        rf_fir_plan* plan = nullptr;
        RF_CHECK(rf_fir_plan_create(static_cast<int>(config.impl),
                                    reinterpret_cast<const float*>(taps_->device()),
                                    static_cast<int>(taps.size()), 0, compute_.s, &plan));
        plan_.reset(plan);
        const bool zero_copy = config.memory == StreamMemory::ZeroCopy;
        if (zero_copy) {
            capture_ = std::make_unique<Memory>(Kind::Mapped, n);
        } else {
            history_ = std::make_unique<Memory>(Kind::Device, halo_);
            if (config.memory == StreamMemory::Registered) {
                capture_ = std::make_unique<Memory>(Kind::Pinned, n);
                result_ = std::make_unique<Memory>(Kind::Pinned, n);
            }
        }
        if (capture_ && n > 0) {
            std::memcpy(capture_->host(), x.data(), n * sizeof(cf32));
        }
        for (int k = 0; k < config.depth; ++k) {
            slots_.push_back(std::make_unique<Slot>());
            Slot& s = *slots_.back();
            s.out = std::make_unique<Memory>(zero_copy ? Kind::Mapped : Kind::Device, window);
            if (!zero_copy) {
                s.window = std::make_unique<Memory>(Kind::Device, window);
            }
            if (config.memory == StreamMemory::Copy) {
                s.stage_in = std::make_unique<Memory>(Kind::Pinned, chunk);
                s.stage_out = std::make_unique<Memory>(Kind::Pinned, chunk);
                std::memset(s.stage_in->host(), 0, chunk * sizeof(cf32));
                std::memset(s.stage_out->host(), 0, chunk * sizeof(cf32));
            }
            if (zero_copy) {
                std::memset(s.out->host(), 0, window * sizeof(cf32));
            }
        }
        if (config.memory == StreamMemory::Registered && n > 0) {
            std::memset(result_->host(), 0, n * sizeof(cf32));
        }
        if (config.memory != StreamMemory::Registered) {
            pool_ = std::make_unique<CopyPool>(copy_threads());
        }
    }

    // Nothing may be freed while a stage still uses it.
    ~Pipeline() { (void)rf_device_synchronize(); }

    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    // Runs every chunk through the pipeline and waits for the last one; returns the chunk count.
    int run()
    {
        const std::size_t n = x_.size();
        const std::size_t chunk = static_cast<std::size_t>(config_.chunk);
        int chunks = 0;
        for (std::size_t start = 0; start < n; start += chunk, ++chunks) {
            Slot& s = *slots_[static_cast<std::size_t>(chunks) % slots_.size()];
            retire(s);
            s.start = start;
            s.count = std::min(chunk, n - start);
            s.hist = std::min(halo_, start);
            enqueue(s);
        }
        for (auto& s : slots_) {
            retire(*s);
        }
        return chunks;
    }

    // Registered mode: the result array is read after the clock stops.
    void collect()
    {
        if (result_ && !y_.empty()) {
            std::memcpy(y_.data(), result_->host(), y_.size() * sizeof(cf32));
        }
    }

private:
    void enqueue(Slot& s)
    {
        const std::size_t bytes = s.count * sizeof(cf32);
        const int win = static_cast<int>(s.hist + s.count);
        if (config_.memory == StreamMemory::ZeroCopy) {
            // the filter reads the mapped capture in place; the window starts hist samples early
            RF_CHECK(rf_fir_plan_run(plan_.get(), as_rf(capture_->device() + (s.start - s.hist)),
                                     as_rf(s.out->device()), win, compute_.s));
            RF_CHECK(rf_event_record(s.filtered.e, compute_.s));
            RF_CHECK(rf_stream_wait_event(download_.s, s.filtered.e));
            RF_CHECK(rf_event_record(s.downloaded.e, download_.s));
            s.busy = true;
            return;
        }
        const cf32* src = nullptr;
        if (config_.memory == StreamMemory::Copy) {
            pool_->copy(s.stage_in->host(), x_.data() + s.start, bytes);
            src = s.stage_in->host();
        } else {
            src = capture_->host() + s.start;
        }
        cf32* window = s.window->device();
        RF_CHECK(rf_memcpy_h2d(window + s.hist, src, bytes, upload_.s));
        RF_CHECK(rf_event_record(s.uploaded.e, upload_.s));

        RF_CHECK(rf_stream_wait_event(compute_.s, s.uploaded.e));
        if (s.hist > 0) {
            RF_CHECK(rf_memcpy_d2d(window, history_->device(), s.hist * sizeof(cf32),
                                   compute_.s));
        }
        RF_CHECK(rf_fir_plan_run(plan_.get(), as_rf(window), as_rf(s.out->device()), win, compute_.s));
        const std::size_t keep = std::min(halo_, s.start + s.count);
        if (keep > 0) {
            RF_CHECK(rf_memcpy_d2d(history_->device(), window + (s.hist + s.count - keep),
                                   keep * sizeof(cf32), compute_.s));
        }
        RF_CHECK(rf_event_record(s.filtered.e, compute_.s));

        RF_CHECK(rf_stream_wait_event(download_.s, s.filtered.e));
        cf32* dst = config_.memory == StreamMemory::Copy ? s.stage_out->host()
                                                          : result_->host() + s.start;
        RF_CHECK(rf_memcpy_d2h(dst, s.out->device() + s.hist, bytes, download_.s));
        RF_CHECK(rf_event_record(s.downloaded.e, download_.s));
        s.busy = true;
    }

    // Waits for the slot's chunk and moves its outputs where they belong.
    void retire(Slot& s)
    {
        if (!s.busy) {
            return;
        }
        RF_CHECK(rf_event_synchronize(s.downloaded.e));
        const std::size_t bytes = s.count * sizeof(cf32);
        if (config_.memory == StreamMemory::Copy) {
            pool_->copy(y_.data() + s.start, s.stage_out->host(), bytes);
        } else if (config_.memory == StreamMemory::ZeroCopy) {
            pool_->copy(y_.data() + s.start, s.out->host() + s.hist, bytes);
        }
        s.busy = false;
    }

    const std::vector<cf32>& x_;
    std::vector<cf32>& y_;
    StreamConfig config_;
    std::size_t halo_;
    StreamHandle upload_, compute_, download_;
    std::unique_ptr<Memory> taps_;
    std::unique_ptr<Memory> history_;
    std::unique_ptr<Memory> capture_;  // registered: pinned copy; zero-copy: mapped copy
    std::unique_ptr<Memory> result_;   // registered: pinned result
    PlanHandle plan_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::unique_ptr<CopyPool> pool_;  // copy and zero-copy modes
};

}  // namespace

bool parse_stream_memory(const std::string& name, StreamMemory& memory)
{
    static const std::pair<const char*, StreamMemory> kModes[] = {
        {"copy", StreamMemory::Copy},
        {"registered", StreamMemory::Registered},
        {"zero-copy", StreamMemory::ZeroCopy},
    };
    for (const auto& m : kModes) {
        if (name == m.first) {
            memory = m.second;
            return true;
        }
    }
    return false;
}

std::vector<cf32> fir_stream(const std::vector<cf32>& x, const std::vector<float>& taps,
                             const StreamConfig& config, StreamTiming& timing)
{
    if (taps.empty()) {
        throw std::runtime_error("stream: no taps");
    }
    if (config.chunk < 1 || config.depth < 1) {
        throw std::runtime_error("stream: chunk and depth must be at least 1");
    }
    std::vector<cf32> y(x.size());
    Pipeline pipeline(x, taps, config, y);
    const auto t0 = std::chrono::steady_clock::now();
    timing.chunks = pipeline.run();
    const auto t1 = std::chrono::steady_clock::now();
    timing.total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    pipeline.collect();
    return y;
}
