// Streaming FIR over a ring of pinned buffers and HIP streams (stream.hpp).

#include "stream.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace {

// Page-locked host buffer; mapped into the device's address space on request.
class Pinned {
public:
    Pinned(std::size_t bytes, bool mapped)
    {
        if (bytes == 0) {
            return;
        }
        if (mapped) {
            RF_CHECK(rf_host_alloc_mapped(&host_, &device_, bytes));
        } else {
            RF_CHECK(rf_host_alloc(&host_, bytes));
        }
    }
    ~Pinned() { rf_host_free(host_); }
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    cf32* host() const { return static_cast<cf32*>(host_); }
    cf32* device() const { return static_cast<cf32*>(device_); }

private:
    void* host_ = nullptr;
    void* device_ = nullptr;
};

class Stream {
public:
    Stream() { RF_CHECK(rf_stream_create(&s_)); }
    ~Stream() { rf_stream_destroy(s_); }
    Stream(const Stream&) = delete;
    Stream& operator=(const Stream&) = delete;
    rf_stream get() const { return s_; }

private:
    rf_stream s_ = nullptr;
};

class Event {
public:
    Event() { RF_CHECK(rf_event_create(&e_)); }
    ~Event() { rf_event_destroy(e_); }
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    void* get() const { return e_; }

private:
    void* e_ = nullptr;
};

struct Slot {
    Slot(std::size_t len, StreamMemory memory)
        : in(memory == StreamMemory::Registered ? 0 : len * sizeof(cf32),
             memory == StreamMemory::ZeroCopy),
          out(memory == StreamMemory::Registered ? 0 : len * sizeof(cf32),
              memory == StreamMemory::ZeroCopy),
          dev_in(memory == StreamMemory::ZeroCopy ? 0 : len),
          dev_out(memory == StreamMemory::ZeroCopy ? 0 : len)
    {
    }
    Pinned in;
    Pinned out;
    DeviceBuffer<cf32> dev_in;
    DeviceBuffer<cf32> dev_out;
    Stream stream;
    Event done;
    std::size_t start = 0;   // first output of the chunk in flight
    std::size_t count = 0;   // its outputs
    std::size_t offset = 0;  // where they sit in the kernel's output
    bool busy = false;
};

}  // namespace

bool parse_stream_memory(const std::string& name, StreamMemory& memory)
{
    if (name == "copy") {
        memory = StreamMemory::Copy;
    } else if (name == "registered") {
        memory = StreamMemory::Registered;
    } else if (name == "zero-copy") {
        memory = StreamMemory::ZeroCopy;
    } else {
        return false;
    }
    return true;
}

std::vector<cf32> fir_stream(const std::vector<cf32>& x, const std::vector<float>& taps,
                             const StreamConfig& config, StreamTiming& timing)
{
    if (config.chunk < 1 || config.depth < 1) {
        throw std::runtime_error("stream: chunk and depth must be at least 1");
    }
    if (taps.empty()) {
        throw std::runtime_error("stream: need at least one tap");
    }
    const std::size_t n = x.size();
    const std::size_t halo = taps.size() - 1;
    const std::size_t chunk = static_cast<std::size_t>(config.chunk);
    const int num_taps = static_cast<int>(taps.size());
    const bool whole = config.memory == StreamMemory::Registered;
    std::vector<cf32> y(n);

    // Setup, outside the clock.
    std::vector<std::unique_ptr<Slot>> slots;
    for (int s = 0; s < config.depth; ++s) {
        slots.push_back(std::make_unique<Slot>(halo + chunk, config.memory));
    }
    Pinned x_pinned(whole ? n * sizeof(cf32) : 0, false);
    Pinned y_pinned(whole ? n * sizeof(cf32) : 0, false);
    if (whole && n > 0) {
        std::memcpy(x_pinned.host(), x.data(), n * sizeof(cf32));
    }
    DeviceBuffer<float> d_taps(taps.size());
    d_taps.upload(taps.data(), taps.size());
    fir_load_const_taps(d_taps.get(), num_taps, nullptr);
    // The filter is the same for every chunk: prepared once (for fft_ola its spectrum too), so
    // each chunk launches only the filtering kernel.
    const FirPlan plan(config.impl, d_taps.get(), num_taps);
    RF_CHECK(rf_device_synchronize());

    const auto retire = [&](Slot& slot) {
        if (!slot.busy) {
            return;
        }
        RF_CHECK(rf_event_synchronize(slot.done.get()));
        if (!whole) {
            std::memcpy(y.data() + slot.start, slot.out.host() + slot.offset,
                        slot.count * sizeof(cf32));
        }
        slot.busy = false;
    };

    const auto begin = std::chrono::steady_clock::now();
    int chunks = 0;
    for (std::size_t start = 0; start < n; start += chunk, ++chunks) {
        Slot& slot = *slots[static_cast<std::size_t>(chunks % config.depth)];
        retire(slot);
        const std::size_t end = std::min(n, start + chunk);
        const std::size_t in_start = start > halo ? start - halo : 0;
        const std::size_t in_len = end - in_start;
        if (!whole) {
            std::memcpy(slot.in.host(), x.data() + in_start, in_len * sizeof(cf32));
        }
        slot.start = start;
        slot.count = end - start;
        slot.offset = start - in_start;
        const rf_stream s = slot.stream.get();
        if (config.memory == StreamMemory::ZeroCopy) {
            plan.run(slot.in.device(), slot.out.device(), static_cast<int>(in_len), s);
        } else {
            const cf32* source = whole ? x_pinned.host() + in_start : slot.in.host();
            RF_CHECK(rf_memcpy_h2d(slot.dev_in.get(), source, in_len * sizeof(cf32), s));
            plan.run(slot.dev_in.get(), slot.dev_out.get(), static_cast<int>(in_len), s);
            cf32* target = whole ? y_pinned.host() + start : slot.out.host() + slot.offset;
            RF_CHECK(rf_memcpy_d2h(target, slot.dev_out.get() + slot.offset,
                                   slot.count * sizeof(cf32), s));
        }
        RF_CHECK(rf_event_record(slot.done.get(), s));
        slot.busy = true;
    }
    for (int i = 0; i < config.depth; ++i) {
        retire(*slots[static_cast<std::size_t>((chunks + i) % config.depth)]);
    }
    const auto finish = std::chrono::steady_clock::now();
    if (whole && n > 0) {
        std::memcpy(y.data(), y_pinned.host(), n * sizeof(cf32));
    }
    timing.total_ms = std::chrono::duration<double, std::milli>(finish - begin).count();
    timing.chunks = chunks;
    return y;
}
