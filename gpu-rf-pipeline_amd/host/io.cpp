// Sample files: raw arrays and SigMF recordings (formats in io.hpp).
//
// A recording is resolved in two steps: its metadata (SigMF .sigmf-meta, or the optional
// NAME.json sidecar of a raw file) is scanned once into a flat table of scalar fields, then the
// data file is decoded by the sample encoding the metadata names. Files are read whole with
// one stat and one read.

#include "io.hpp"

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <map>
#include <stdexcept>
#include <string_view>

namespace fs = std::filesystem;

namespace {

[[noreturn]] void file_error(const std::string& path, const std::string& what)
{
    throw std::runtime_error(path + ": " + what);
}

// The bytes of a file.
std::vector<unsigned char> read_bytes(const std::string& path)
{
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (f == nullptr) {
        file_error(path, std::string("cannot open (") + std::strerror(errno) + ")");
    }
    std::vector<unsigned char> bytes;
    std::error_code ec;
    const auto size = fs::file_size(path, ec);
    if (!ec) {
        bytes.resize(static_cast<std::size_t>(size));
    }
    const std::size_t got = bytes.empty() ? 0 : std::fread(bytes.data(), 1, bytes.size(), f);
    const bool failed = std::ferror(f) != 0 || got != bytes.size();
    std::fclose(f);
    if (failed) {
        file_error(path, "read failed");
    }
    return bytes;
}

std::string read_text(const std::string& path)
{
    const std::vector<unsigned char> b = read_bytes(path);
    return std::string(b.begin(), b.end());
}


// Metadata documents are small JSON files. JsonScalars walks one with a recursive-descent
// scanner and keeps every `"key": scalar` pair it meets, at any depth (SigMF puts its fields
// under "global"); the first occurrence of a key wins. Strings keep their text without the
// quotes (escapes are kept verbatim); numbers and literals keep their spelling.
class JsonScalars {
public:
    JsonScalars(std::string_view text, const std::string& origin) : s_(text), origin_(origin)
    {
        skip_space();
        value("");
        skip_space();
        if (i_ != s_.size()) {
            bad("trailing text");
        }
    }

    const std::string* find(const std::string& key) const
    {
        const auto it = fields_.find(key);
        return it == fields_.end() ? nullptr : &it->second;
    }

private:
    [[noreturn]] void bad(const char* what) const
    {
        file_error(origin_, std::string("malformed JSON metadata (") + what + " at byte "
                                + std::to_string(i_) + ")");
    }

    void skip_space()
    {
        while (i_ < s_.size() && std::strchr(" \t\r\n", s_[i_]) != nullptr && s_[i_] != '\0') {
            ++i_;
        }
    }

    std::string string_token()
    {
        ++i_;  // opening quote
        const std::size_t begin = i_;
        while (i_ < s_.size() && s_[i_] != '"') {
            i_ += s_[i_] == '\\' ? 2 : 1;
        }
        if (i_ >= s_.size()) {
            bad("unterminated string");
        }
        return std::string(s_.substr(begin, i_++ - begin));
    }

    // A value; `key` is the member name it belongs to ("" inside arrays and at the top).
    void value(const std::string& key)
    {
        if (i_ >= s_.size()) {
            bad("missing value");
        }
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            skip_space();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return;
            }
            for (;;) {
                skip_space();
                if (i_ >= s_.size() || s_[i_] != '"') {
                    bad("expected a member name");
                }
                const std::string name = string_token();
                skip_space();
                if (i_ >= s_.size() || s_[i_] != ':') {
                    bad("expected ':'");
                }
                ++i_;
                skip_space();
                value(name);
                skip_space();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return;
                }
                bad("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i_;
            skip_space();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return;
            }
            for (;;) {
                skip_space();
                value("");
                skip_space();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    return;
                }
                bad("expected ',' or ']'");
            }
        }
        std::string scalar;
        if (c == '"') {
            scalar = string_token();
        } else {
            const std::size_t begin = i_;
            while (i_ < s_.size() && std::strchr(",}] \t\r\n", s_[i_]) == nullptr) {
                ++i_;
            }
            if (i_ == begin) {
                bad("empty value");
            }
            scalar = std::string(s_.substr(begin, i_ - begin));
        }
        if (!key.empty()) {
            fields_.emplace(key, std::move(scalar));
        }
    }

    std::string_view s_;
    std::string origin_;
    std::size_t i_ = 0;
    std::map<std::string, std::string> fields_;
};

// How one complex sample is stored: two components of `bytes` each, integers scaled by `scale`.
struct Encoding {
    const char* name;
    std::size_t bytes;
    bool integer;
    float scale;
};

constexpr Encoding kEncodings[] = {
    {"cf32_le", 4, false, 1.0f},
    {"ci16_le", 2, true, 1.0f / 32768.0f},
    {"ci8", 1, true, 1.0f / 128.0f},
    {"ci8_le", 1, true, 1.0f / 128.0f},
};

std::vector<cf32> decode(const std::vector<unsigned char>& bytes, const Encoding& enc,
                         const std::string& path)
{
    const std::size_t per_sample = 2 * enc.bytes;
    if (bytes.size() % per_sample != 0) {
        file_error(path, std::to_string(bytes.size()) + " bytes is not a whole number of "
                             + enc.name + " samples (" + std::to_string(per_sample)
                             + " bytes each)");
    }
    std::vector<cf32> out(bytes.size() / per_sample);
    if (!enc.integer) {
        if (!out.empty()) {
            std::memcpy(out.data(), bytes.data(), bytes.size());
        }
        return out;
    }
    float* iq = reinterpret_cast<float*>(out.data());
    for (std::size_t k = 0; k < 2 * out.size(); ++k) {
        const unsigned char* p = bytes.data() + k * enc.bytes;
        const int v = enc.bytes == 1 ? static_cast<int>(static_cast<std::int8_t>(p[0]))
                                     : static_cast<int>(static_cast<std::int16_t>(
                                           static_cast<std::uint16_t>(p[0] | (p[1] << 8))));
        iq[k] = static_cast<float>(v) * enc.scale;
    }
    return out;
}

bool ends_with(const std::string& s, std::string_view tail)
{
    return s.size() >= tail.size() && std::string_view(s).substr(s.size() - tail.size()) == tail;
}

// NAME of NAME.sigmf-meta / NAME.sigmf-data.
std::string sigmf_name(const std::string& path)
{
    return path.substr(0, path.rfind(".sigmf-"));
}

std::vector<cf32> read_recording(const std::string& path)
{
    const std::string name = sigmf_name(path);
    const std::string meta_path = name + ".sigmf-meta";
    const JsonScalars meta(read_text(meta_path), meta_path);
    if (const std::string* channels = meta.find("core:num_channels")) {
        if (std::strtol(channels->c_str(), nullptr, 10) != 1) {
            file_error(meta_path, "recording has " + *channels
                                      + " channels; only single-channel recordings are read");
        }
    }
    const std::string* type = meta.find("core:datatype");
    const std::string datatype = type == nullptr ? "(none)" : *type;
    for (const Encoding& enc : kEncodings) {
        if (datatype == enc.name) {
            return decode(read_bytes(name + ".sigmf-data"), enc, name + ".sigmf-data");
        }
    }
    file_error(meta_path, "core:datatype '" + datatype
                              + "' is not supported (cf32_le, ci16_le, ci8)");
}

}  // namespace

bool is_sigmf(const std::string& path)
{
    return ends_with(path, ".sigmf-meta") || ends_with(path, ".sigmf-data");
}

double sample_rate(const std::string& path)
{
    const bool sigmf = is_sigmf(path);
    const std::string meta_path =
        sigmf ? sigmf_name(path) + ".sigmf-meta" : fs::path(path).replace_extension(".json").string();
    std::error_code ec;
    if (!fs::exists(meta_path, ec)) {
        return 0.0;
    }
    const JsonScalars meta(read_text(meta_path), meta_path);
    const std::string* rate = meta.find(sigmf ? "core:sample_rate" : "sample_rate");
    return rate == nullptr ? 0.0 : std::strtod(rate->c_str(), nullptr);
}

void write_raw(const std::string& path, const void* data, std::size_t bytes)
{
    const fs::path dir = fs::path(path).parent_path();
    std::error_code ec;
    if (!dir.empty() && !fs::exists(dir, ec)) {
        fs::create_directories(dir, ec);
        if (ec) {
            file_error(dir.string(), "cannot create directory (" + ec.message() + ")");
        }
    }
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (f == nullptr) {
        file_error(path, std::string("cannot open for writing (") + std::strerror(errno) + ")");
    }
    const std::size_t put = bytes == 0 ? 0 : std::fwrite(data, 1, bytes, f);
    const bool failed = put != bytes || std::fclose(f) != 0;
    if (failed) {
        file_error(path, "write failed");
    }
}

std::vector<cf32> read_cf32(const std::string& path)
{
    return is_sigmf(path) ? read_recording(path) : decode(read_bytes(path), kEncodings[0], path);
}

std::vector<float> read_f32(const std::string& path)
{
    const std::vector<unsigned char> bytes = read_bytes(path);
    if (bytes.size() % sizeof(float) != 0) {
        file_error(path, std::to_string(bytes.size())
                             + " bytes is not a whole number of float32 values");
    }
    std::vector<float> out(bytes.size() / sizeof(float));
    if (!out.empty()) {
        std::memcpy(out.data(), bytes.data(), bytes.size());
    }
    return out;
}

void write_cf32(const std::string& path, const std::vector<cf32>& samples)
{
    write_raw(path, samples.data(), samples.size() * sizeof(cf32));
}

void write_f32(const std::string& path, const std::vector<float>& values)
{
    write_raw(path, values.data(), values.size() * sizeof(float));
}
