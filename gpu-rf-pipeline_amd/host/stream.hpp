#pragma once

// Streaming FIR (reference: src/stream.hpp:1-57, src/stream.cu:135-222): the capture is cut
// into chunks that go round a ring of `depth` slots, each with pinned host buffers, device
// buffers and its own HIP stream, so the copy engines and the CUs work on different chunks
// at once. Every chunk after the first is uploaded with num_taps - 1 samples of history and
// only its own outputs come back, so for the direct FIR variants the result is bit-identical
// to one rf_fir over the whole capture. fft_ola places its transform blocks relative to each
// chunk, so its streamed result equals the one-shot result to rounding, not bit for bit.
//
// Memory modes:
//   copy        host memcpy into and out of the pinned ring, async copies to the device;
//   registered  the capture and the result live in pinned memory for the whole run (the
//               receiver writes into pinned memory), chunks are copied straight from and to it;
//               the pinned arrays are allocated and filled before the clock starts;
//   zero-copy   mapped pinned buffers, the kernel reads and writes host memory directly.

#include "device_ops.hpp"
#include "types.hpp"

#include <string>
#include <vector>

enum class StreamMemory { Copy, Registered, ZeroCopy };

struct StreamConfig {
    FirImpl impl = FirImpl::Tiled;
    int chunk = 1 << 20;
    int depth = 3;
    StreamMemory memory = StreamMemory::Copy;
};

bool parse_stream_memory(const std::string& name, StreamMemory& memory);

struct StreamTiming {
    double total_ms = 0.0;  // wall clock from the first chunk to the last output on the host
    int chunks = 0;
};

std::vector<cf32> fir_stream(const std::vector<cf32>& x, const std::vector<float>& taps,
                             const StreamConfig& config, StreamTiming& timing);
