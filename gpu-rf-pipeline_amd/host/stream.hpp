#pragma once

// Streaming FIR (interface of the reference's src/stream.hpp:1-57): a capture filtered chunk by
// chunk the way a receiver delivers it, with copies and filtering of different chunks
// overlapped. The output equals one FIR over the whole capture: bit for bit for the direct
// variants; to rounding for fft_ola and the MFMA variants, whose transform / GEMM blocks sit
// relative to each chunk's window.
//
// `depth` chunks may be in flight at once (a ring of that many slots); `chunk` samples each.
// Memory modes:
//   copy        the host copies each chunk into pinned staging memory, which is copied to the
//               device, and the outputs back out of pinned staging;
//   registered  capture and result live in pinned memory for the whole run (a receiver that
//               writes into pinned memory), chunks are copied straight from and to them; they
//               are filled before and read after the clock;
//   zero-copy   capture and outputs in mapped pinned memory; the filter reads and writes host
//               memory directly, with no copies to or from the device.

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
