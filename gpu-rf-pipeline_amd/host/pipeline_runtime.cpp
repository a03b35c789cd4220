// sigflow::Runtime (pipeline_runtime.hpp). Reference behaviour: src/pipeline_runtime.cpp:62-184
// (flags, slot sizing, taps cache, plan ownership, counted first pass, JSON report).

#include "pipeline_runtime.hpp"

#include "io.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sigflow {

namespace {

int int_flag(int argc, char** argv, const char* name, int fallback, int minimum)
{
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], name) != 0) {
            continue;
        }
        if (i + 1 >= argc) {
            throw std::runtime_error(std::string(name) + " needs a value");
        }
        char* end = nullptr;
        const long v = std::strtol(argv[i + 1], &end, 10);
        if (end == argv[i + 1] || *end != '\0' || v < minimum || v > 1000000) {
            throw std::runtime_error(std::string(name) + " must be an integer of at least "
                                     + std::to_string(minimum));
        }
        return static_cast<int>(v);
    }
    return fallback;
}

bool has_flag(int argc, char** argv, const char* name)
{
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], name) == 0) {
            return true;
        }
    }
    return false;
}

std::string json_text(const std::string& s)
{
    std::string out = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') {
            out += '\\';
        }
        out += static_cast<unsigned char>(c) < 0x20 ? ' ' : c;
    }
    return out + "\"";
}

}  // namespace

template <class T>
HostArray<T>::HostArray(std::size_t count) : count_(count)
{
    if (count_ > 0) {
        void* p = nullptr;
        RF_CHECK(rf_host_alloc(&p, bytes()));
        std::memset(p, 0, bytes());
        data_ = static_cast<T*>(p);
    }
}

template <class T>
HostArray<T>::~HostArray()
{
    if (data_ != nullptr) {
        (void)rf_host_free(data_);
    }
}

template class HostArray<cf32>;
template class HostArray<float>;

Runtime::Runtime(const char* program, const char* kernels, bool fused, bool reuse, int argc,
                 char** argv)
    : program_(program), kernels_(kernels), fused_(fused), reuse_(reuse)
{
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--bench" || a == "--warmup") {
            ++i;
        } else if (a != "--graph") {
            throw std::runtime_error("unknown argument '" + a
                                     + "' (usage: PROGRAM [--bench N] [--warmup W] [--graph])");
        }
    }
    config_.runs = int_flag(argc, argv, "--bench", 1, 1);
    config_.warmup = int_flag(argc, argv, "--warmup", 0, 0);
    graph_ = has_flag(argc, argv, "--graph");
    RF_CHECK(rf_stream_create(&stream_));
    config_.stream = stream_;
}

Runtime::~Runtime()
{
    if (stream_ != nullptr) {
        (void)rf_stream_synchronize(stream_);
    }
    if (graph_exec_ != nullptr) {
        (void)rf_graph_destroy(graph_exec_);
    }
    xcorrs_.clear();
    plans_.clear();
    for (auto& t : taps_) {
        (void)rf_free(t.second.first);
    }
    for (void* p : slots_) {
        if (p != nullptr) {
            (void)rf_free(p);
        }
    }
    if (stream_ != nullptr) {
        (void)rf_stream_destroy(stream_);
    }
}

const HostArray<cf32>& Runtime::read_capture(const std::string& path)
{
    const std::vector<cf32> samples = ::read_cf32(path);
    captures_.push_back(std::make_unique<HostArray<cf32>>(samples.size()));
    if (!samples.empty()) {
        std::memcpy(captures_.back()->data(), samples.data(), samples.size() * sizeof(cf32));
    }
    return *captures_.back();
}

HostArray<cf32>& Runtime::host_cf32(std::size_t count)
{
    host_cf32_.push_back(std::make_unique<HostArray<cf32>>(count));
    return *host_cf32_.back();
}

HostArray<float>& Runtime::host_f32(std::size_t count)
{
    host_f32_.push_back(std::make_unique<HostArray<float>>(count));
    return *host_f32_.back();
}

void Runtime::need(int slot, std::size_t bytes)
{
    if (slot < 0) {
        throw std::runtime_error("need: negative slot");
    }
    const std::size_t i = static_cast<std::size_t>(slot);
    if (slot_bytes_.size() <= i) {
        slot_bytes_.resize(i + 1, 0);
    }
    slot_bytes_[i] = std::max(slot_bytes_[i], bytes);
}

void Runtime::allocate_slots()
{
    if (!slots_.empty()) {
        throw std::runtime_error("allocate_slots: called twice");
    }
    for (std::size_t bytes : slot_bytes_) {
        void* p = nullptr;
        if (bytes > 0) {
            RF_CHECK(rf_malloc(&p, bytes));
        }
        slots_.push_back(p);
    }
}

Taps Runtime::taps(const std::string& path)
{
    auto it = taps_.find(path);
    if (it == taps_.end()) {
        const std::vector<float> host = ::read_f32(path);
        if (host.empty()) {
            throw std::runtime_error(path + ": no taps");
        }
        void* p = nullptr;
        RF_CHECK(rf_malloc(&p, host.size() * sizeof(float)));
        RF_CHECK(rf_memcpy_h2d(p, host.data(), host.size() * sizeof(float), stream_));
        RF_CHECK(rf_stream_synchronize(stream_));
        taps_bytes_ += host.size() * sizeof(float);
        it = taps_.emplace(path, std::make_pair(p, samples(host.size()))).first;
    }
    return Taps{static_cast<const float*>(it->second.first), it->second.second};
}

const rf_fft_plan* Runtime::fft_plan(long long nfft, std::size_t frames)
{
    if (nfft < 1 || nfft > INT_MAX) {
        throw std::runtime_error("fft_plan: bad length " + std::to_string(nfft));
    }
    plans_.push_back(std::make_unique<FftPlan>(static_cast<int>(nfft), samples(frames)));
    std::size_t bytes = 0;
    RF_CHECK(rf_fft_plan_bytes(plans_.back()->get(), &bytes));
    fft_work_bytes_ += bytes;
    return plans_.back()->get();
}

const XcorrFft& Runtime::xcorr_fft(std::size_t num_a, std::size_t num_b, bool phat)
{
    xcorrs_.push_back(std::make_unique<XcorrFft>(samples(num_a), samples(num_b), phat));
    fft_work_bytes_ += xcorrs_.back()->bytes();
    return *xcorrs_.back();
}

void Runtime::upload(int slot, const HostArray<cf32>& host)
{
    RF_CHECK(rf_memcpy_h2d(this->slot<cf32>(slot), host.data(), host.bytes(), stream_));
}

template <class T>
void Runtime::download(HostArray<T>& host, int slot)
{
    RF_CHECK(rf_memcpy_d2h(host.data(), this->slot<T>(slot), host.bytes(), stream_));
}

template void Runtime::download<cf32>(HostArray<cf32>&, int);
template void Runtime::download<float>(HostArray<float>&, int);

void Runtime::zero_tail(int slot, std::size_t from, std::size_t to)
{
    if (to > from) {
        RF_CHECK(rf_memset(this->slot<unsigned char>(slot) + from, 0, to - from, stream_));
    }
}

void Runtime::run(const std::function<void()>& upload, const std::function<void()>& compute,
                  const std::function<void()>& download)
{
    RF_CHECK(rf_stream_synchronize(stream_));
    std::function<void()> timed = compute;
    if (graph_) {
        // One eager pass first (counted, and any first-call set-up of the library happens
        // outside the recording), then the compute phase recorded once and replayed per run.
        upload();
        counting_ = true;
        compute();
        counting_ = false;
        download();
        RF_CHECK(rf_stream_synchronize(stream_));
        RF_CHECK(rf_graph_begin_capture(stream_));
        try {
            compute();
        } catch (...) {
            void* discard = nullptr;
            (void)rf_graph_end_capture(stream_, &discard);
            if (discard != nullptr) {
                (void)rf_graph_destroy(discard);
            }
            throw;
        }
        RF_CHECK(rf_graph_end_capture(stream_, &graph_exec_));
        timed = [this] { RF_CHECK(rf_graph_launch(graph_exec_, stream_)); };
    }
    bool first = counts_.empty();
    Bench bench(config_);
    bench.run(upload,
              [&] {
                  counting_ = first;
                  first = false;
                  timed();
                  counting_ = false;
              },
              download);
    RF_CHECK(rf_stream_synchronize(stream_));
    result_ = bench.result();
}

void Runtime::write_cf32(const std::string& path, const HostArray<cf32>& data) const
{
    write_raw(path, data.data(), data.bytes());
}

void Runtime::write_f32(const std::string& path, const HostArray<float>& data) const
{
    write_raw(path, data.data(), data.bytes());
}

void Runtime::report() const
{
    std::size_t slot_total = 0;
    for (std::size_t b : slot_bytes_) {
        slot_total += b;
    }
    int calls = 0;
    std::string by_kind = "{";
    for (const auto& [kind, n] : counts_) {
        by_kind += (by_kind.size() > 1 ? ", " : "") + json_text(kind) + ": " + std::to_string(n);
        calls += n;
    }
    by_kind += "}";
    char gpu[256] = {0};
    if (rf_device_name(gpu, static_cast<int>(sizeof(gpu))) != RF_OK) {
        std::snprintf(gpu, sizeof(gpu), "unknown");
    }
    const BenchResult& r = result_;
    std::printf("{\"program\": %s, \"kernels\": %s, \"fused\": %s, \"reuse\": %s, \"graph\": %s, "
                "\"gpu\": %s, \"calls\": %d, \"calls_by_kind\": %s, \"slots\": %zu, "
                "\"slot_bytes\": %zu, \"taps_bytes\": %zu, \"fft_work_bytes\": %zu, "
                "\"allocated_bytes\": %zu, \"runs\": %d, \"warmup\": %d, "
                "\"h2d_ms\": %.6f, \"h2d_std_ms\": %.6f, \"kernel_ms\": %.6f, "
                "\"kernel_std_ms\": %.6f, \"d2h_ms\": %.6f, \"d2h_std_ms\": %.6f}\n",
                json_text(program_).c_str(), json_text(kernels_).c_str(),
                fused_ ? "true" : "false", reuse_ ? "true" : "false", graph_ ? "true" : "false",
                json_text(gpu).c_str(), calls, by_kind.c_str(), slot_bytes_.size(), slot_total,
                taps_bytes_, fft_work_bytes_, slot_total + taps_bytes_ + fft_work_bytes_,
                config_.runs, config_.warmup, r.h2d.mean_ms, r.h2d.std_ms, r.kernel.mean_ms,
                r.kernel.std_ms, r.d2h.mean_ms, r.d2h.std_ms);
    std::fflush(stdout);
}

}  // namespace sigflow
