// Per-stream workspace of the one-shot calls that derive a small table from their arguments
// before the main kernel (the overlap-save tap spectrum, the split taps of the matrix-core FIR).
// Internal to the library.
#pragma once

#include "rf_common.hpp"

#include <cstddef>

namespace rf {

// One block per (device, stream): covers the 4096-point tap spectrum (32 KiB) and the split
// taps of a 2049-tap matrix-core filter (68 368 bytes).
constexpr std::size_t kStreamWorkspaceBytes = 72 * 1024;

// The block of `stream` on the current device. The first call on a stream allocates it (with
// the calling thread's capture mode relaxed for that one allocation, so it also succeeds as the
// first call inside a stream capture); every later call on the stream returns the same block
// and allocates nothing. Work on one stream runs in order, so calls on it never share the block
// at the same time. A graph captured on a stream keeps that stream's block: replays on the
// capture stream are ordered with eager calls on it.
void* stream_workspace(hipStream_t stream);

// Frees the block of `stream` (after the stream's work has finished), if it has one.
void release_stream_workspace(hipStream_t stream);

// True while `stream` records a graph (or when the query itself fails, e.g. the null stream
// during another stream's global capture).
bool stream_capturing(hipStream_t stream);

}  // namespace rf
