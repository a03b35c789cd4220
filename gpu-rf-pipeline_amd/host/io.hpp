#pragma once

// Sample files (formats of the reference's src/io.hpp:1-30): raw little-endian .cf32
// (interleaved I/Q float32) and .f32 (float32 array), and SigMF recordings (cf32_le,
// ci16_le, ci8; integers scaled by 1/32768 and 1/128). Every function throws
// std::runtime_error naming the file on failure. Little-endian host assumed.

#include "types.hpp"

#include <string>
#include <vector>

std::vector<cf32> read_cf32(const std::string& path);
void write_cf32(const std::string& path, const std::vector<cf32>& samples);
std::vector<float> read_f32(const std::string& path);
void write_f32(const std::string& path, const std::vector<float>& values);
// `bytes` bytes to `path`, creating its parent directories.
void write_raw(const std::string& path, const void* data, std::size_t bytes);

// True for NAME.sigmf-meta or NAME.sigmf-data.
bool is_sigmf(const std::string& path);

// "core:sample_rate" of a SigMF recording, "sample_rate" of the PATH.json sidecar of a raw
// file, or 0 when neither exists.
double sample_rate(const std::string& path);
