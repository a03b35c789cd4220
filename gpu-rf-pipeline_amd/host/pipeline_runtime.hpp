#pragma once

// Runtime of the programs sigflowc writes with --emit=hip (reference: src/pipeline_runtime.hpp,
// the runtime of its --emit=cuda programs). A generated program keeps every value in GPU memory,
// in the buffer slots liveness analysis assigned, and queues all its work on the runtime's one
// stream through host/device_ops.hpp. The runtime does what every program shares:
//
//   setup     read the captures into page-locked host memory, size and allocate the slots,
//             upload the taps, make the FFT and cross-correlation plans; all before timing
//   run       upload / compute / download through Bench (device events on the stream); with
//             --graph the compute phase is recorded once as a hipGraph and replayed
//   report    one JSON object on stdout: library calls by kind, GPU memory, phase times
//
// Command line of a generated program:
//   PROGRAM [--bench N] [--warmup W] [--graph]    default: one run, no warm-up, no graph
//
// Calls (rt.count) are operator calls per pipeline run; an FFT-based call enqueues several
// kernels.

#include "bench.hpp"
#include "device_ops.hpp"
#include "types.hpp"

#include <climits>
#include <cstddef>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sigflow {

// Samples in ceil(n / frame) whole frames: what an FFT of n samples reads and produces.
inline std::size_t padded_to_frames(std::size_t n, long long frame)
{
    const std::size_t f = static_cast<std::size_t>(frame);
    return (n + f - 1) / f * f;
}

// A sample count as the int the operators take; throws above INT_MAX.
inline int samples(std::size_t n)
{
    if (n > static_cast<std::size_t>(INT_MAX)) {
        throw std::runtime_error(std::to_string(n) + " samples: more than one call takes");
    }
    return static_cast<int>(n);
}

// Page-locked host memory for `count` values of T (zero-initialised), so uploads and downloads
// run at full DMA rate and asynchronously on the runtime's stream.
template <class T>
class HostArray {
public:
    explicit HostArray(std::size_t count);
    ~HostArray();
    HostArray(const HostArray&) = delete;
    HostArray& operator=(const HostArray&) = delete;

    T* data() { return data_; }
    const T* data() const { return data_; }
    std::size_t size() const { return count_; }
    std::size_t bytes() const { return count_ * sizeof(T); }
    T& operator[](std::size_t i) { return data_[i]; }
    const T& operator[](std::size_t i) const { return data_[i]; }

private:
    T* data_ = nullptr;
    std::size_t count_ = 0;
};

// A taps file in GPU memory.
struct Taps {
    const float* data = nullptr;
    int count = 0;
};

class Runtime {
public:
    // `program` names the source in the report and `kernels` the variants it was compiled for
    // ("fir=naive xcorr=fft"). Parses the command line and creates the stream.
    Runtime(const char* program, const char* kernels, bool fused, bool reuse, int argc,
            char** argv);
    ~Runtime();
    Runtime(const Runtime&) = delete;
    Runtime& operator=(const Runtime&) = delete;

    rf_stream stream() const { return stream_; }

    // --- setup -----------------------------------------------------------------------------

    const HostArray<cf32>& read_capture(const std::string& path);
    // Page-locked host buffers for downloaded values, owned by the runtime.
    HostArray<cf32>& host_cf32(std::size_t count);
    HostArray<float>& host_f32(std::size_t count);

    // Slot `slot` must hold at least `bytes`; call for every value (an FFT input with its
    // padded length) before allocate_slots().
    void need(int slot, std::size_t bytes);
    void allocate_slots();
    template <class T>
    T* slot(int index) const
    {
        return reinterpret_cast<T*>(slots_.at(static_cast<std::size_t>(index)));
    }

    // Taps from a .f32 file, uploaded once per distinct path.
    Taps taps(const std::string& path);
    // A batched FFT plan (`frames` transforms of `nfft` points), owned by the runtime.
    const rf_fft_plan* fft_plan(long long nfft, std::size_t frames);
    // An FFT cross-correlation (GCC-PHAT with `phat`) for these lengths, owned by the runtime.
    const XcorrFft& xcorr_fft(std::size_t num_a, std::size_t num_b, bool phat = false);

    // --- inside run ------------------------------------------------------------------------

    void upload(int slot, const HostArray<cf32>& host);
    template <class T>
    void download(HostArray<T>& host, int slot);
    // Zeroes bytes [from, to) of a slot (an FFT input's last-frame padding); a fill kernel, so
    // it can be recorded in a graph. Nothing when to <= from.
    void zero_tail(int slot, std::size_t from, std::size_t to);
    // Records that compute() makes one operator call of this kind (first pass only).
    void count(const char* kind)
    {
        if (counting_) {
            ++counts_[kind];
        }
    }

    void run(const std::function<void()>& upload, const std::function<void()>& compute,
             const std::function<void()>& download);

    // --- results ---------------------------------------------------------------------------

    void write_cf32(const std::string& path, const HostArray<cf32>& data) const;
    void write_f32(const std::string& path, const HostArray<float>& data) const;
    void report() const;

private:
    std::string program_;
    std::string kernels_;
    bool fused_;
    bool reuse_;
    bool graph_ = false;
    BenchConfig config_;
    rf_stream stream_ = nullptr;
    std::vector<std::size_t> slot_bytes_;
    std::vector<void*> slots_;
    std::vector<std::unique_ptr<HostArray<cf32>>> captures_;
    std::vector<std::unique_ptr<HostArray<cf32>>> host_cf32_;
    std::vector<std::unique_ptr<HostArray<float>>> host_f32_;
    std::map<std::string, std::pair<void*, int>> taps_;
    std::vector<std::unique_ptr<FftPlan>> plans_;
    std::vector<std::unique_ptr<XcorrFft>> xcorrs_;
    std::size_t taps_bytes_ = 0;
    std::size_t fft_work_bytes_ = 0;
    std::map<std::string, int> counts_;
    bool counting_ = false;
    BenchResult result_;
    void* graph_exec_ = nullptr;  // the recorded compute phase (--graph)
};

}  // namespace sigflow
