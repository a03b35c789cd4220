#include "io.hpp"

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace fs = std::filesystem;

namespace {

std::string slurp(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        throw std::runtime_error("cannot open " + path);
    }
    return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

template <class T>
std::vector<T> load(const std::string& path, const char* unit)
{
    const std::string bytes = slurp(path);
    if (bytes.size() % sizeof(T) != 0) {
        throw std::runtime_error(path + ": " + std::to_string(bytes.size())
                                 + " bytes is not a whole number of " + unit + " ("
                                 + std::to_string(sizeof(T)) + " bytes each)");
    }
    std::vector<T> out(bytes.size() / sizeof(T));
    if (!out.empty()) {
        std::memcpy(out.data(), bytes.data(), bytes.size());
    }
    return out;
}

template <class T>
void store(const std::string& path, const T* data, std::size_t count)
{
    const fs::path parent = fs::path(path).parent_path();
    if (!parent.empty()) {
        std::error_code ec;
        fs::create_directories(parent, ec);
        if (ec) {
            throw std::runtime_error("cannot create directory " + parent.string() + ": "
                                     + ec.message());
        }
    }
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) {
        throw std::runtime_error("cannot open " + path + " for writing");
    }
    out.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
    out.close();
    if (!out) {
        throw std::runtime_error("failed while writing " + path);
    }
}

bool has_suffix(const std::string& s, const std::string& suffix)
{
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

std::string sigmf_stem(const std::string& path)
{
    return path.substr(0, path.size() - 11);  // both suffixes are 11 characters
}

// Raw text of the value of the first "key": in a flat JSON document (a string without its
// quotes, or a number); empty when absent.
std::string json_value(const std::string& doc, const std::string& key)
{
    const std::string quoted = "\"" + key + "\"";
    std::size_t at = doc.find(quoted);
    if (at == std::string::npos) {
        return "";
    }
    at = doc.find(':', at + quoted.size());
    if (at == std::string::npos) {
        return "";
    }
    at = doc.find_first_not_of(" \t\r\n", at + 1);
    if (at == std::string::npos) {
        return "";
    }
    if (doc[at] == '"') {
        const std::size_t end = doc.find('"', at + 1);
        return end == std::string::npos ? "" : doc.substr(at + 1, end - at - 1);
    }
    const std::size_t end = doc.find_first_of(",}] \t\r\n", at);
    return doc.substr(at, end == std::string::npos ? std::string::npos : end - at);
}

template <class I>
std::vector<cf32> integers_to_cf32(const std::string& path, float scale)
{
    const std::vector<I> raw = load<I>(path, "integer components");
    if (raw.size() % 2 != 0) {
        throw std::runtime_error(path + ": odd number of components, not I/Q pairs");
    }
    std::vector<cf32> out(raw.size() / 2);
    for (std::size_t i = 0; i < out.size(); ++i) {
        out[i] = cf32{static_cast<float>(raw[2 * i]) * scale,
                      static_cast<float>(raw[2 * i + 1]) * scale};
    }
    return out;
}

std::vector<cf32> read_sigmf(const std::string& path)
{
    const std::string stem = sigmf_stem(path);
    const std::string meta = slurp(stem + ".sigmf-meta");
    const std::string data = stem + ".sigmf-data";
    const std::string type = json_value(meta, "core:datatype");
    const std::string channels = json_value(meta, "core:num_channels");
    if (!channels.empty() && std::atoi(channels.c_str()) != 1) {
        throw std::runtime_error(stem + ".sigmf-meta: " + channels
                                 + " channels; only single-channel recordings are supported");
    }
    if (type == "cf32_le") {
        return load<cf32>(data, "complex float32 samples");
    }
    if (type == "ci16_le") {
        return integers_to_cf32<std::int16_t>(data, 1.0f / 32768.0f);
    }
    if (type == "ci8" || type == "ci8_le") {
        return integers_to_cf32<std::int8_t>(data, 1.0f / 128.0f);
    }
    throw std::runtime_error(stem + ".sigmf-meta: core:datatype '" + type
                             + "' is not supported (cf32_le, ci16_le, ci8)");
}

}  // namespace

bool is_sigmf(const std::string& path)
{
    return has_suffix(path, ".sigmf-meta") || has_suffix(path, ".sigmf-data");
}

double sample_rate(const std::string& path)
{
    std::string meta;
    std::string key;
    if (is_sigmf(path)) {
        meta = sigmf_stem(path) + ".sigmf-meta";
        key = "core:sample_rate";
    } else {
        meta = fs::path(path).replace_extension(".json").string();
        key = "sample_rate";
        if (!fs::exists(meta)) {
            return 0.0;
        }
    }
    const std::string v = json_value(slurp(meta), key);
    return v.empty() ? 0.0 : std::strtod(v.c_str(), nullptr);
}

std::vector<cf32> read_cf32(const std::string& path)
{
    return is_sigmf(path) ? read_sigmf(path) : load<cf32>(path, "complex float32 samples");
}

std::vector<float> read_f32(const std::string& path)
{
    return load<float>(path, "float32 values");
}

void write_cf32(const std::string& path, const std::vector<cf32>& samples)
{
    store(path, samples.data(), samples.size());
}

void write_f32(const std::string& path, const std::vector<float>& values)
{
    store(path, values.data(), values.size());
}
