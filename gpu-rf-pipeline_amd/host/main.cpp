// rfgpu: command-line front end of the MI355X hot path, with the reference's subcommands,
// flags, JSON reports and exit codes (src/main.cpp:1-549): 0 success, 1 runtime error (file,
// HIP), 2 usage error.
//
//   rfgpu spectrum  --in x.cf32 --out spec.cf32 [--nfft 1024]
//   rfgpu fir       --in x.cf32 --taps h.f32 --out y.cf32 [--impl naive] [--block 256]
//   rfgpu xcorr     --a a.cf32 --b b.cf32 --out r.cf32 [--impl naive]
//   rfgpu magnitude --in x.cf32 --out mag.f32
//   rfgpu channelize --in x.cf32 --taps h.f32 --channels M --out z.cf32
//   rfgpu gcc_phat  --a a.cf32 --b b.cf32 --out r.cf32
//   rfgpu doa       --in ch0.cf32,ch1.cf32,... --spacing METERS [--speed M_PER_S] [--rate HZ]
//   rfgpu stream    --in x.cf32 --taps h.f32 --out y.cf32 [--chunk N] [--depth D] [--memory M]
//   rfgpu info

#include "bench.hpp"
#include "check.hpp"
#include "io.hpp"
#include "ops.hpp"
#include "stream.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <exception>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct UsageError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class Fn>
std::string names_of(const std::vector<Impl<Fn>>& impls, const char* sep)
{
    std::string out;
    for (const auto& i : impls) {
        out += (out.empty() ? "" : sep);
        out += i.name;
    }
    return out;
}

template <class Fn>
void list_impls(const char* op, const std::vector<Impl<Fn>>& impls)
{
    for (std::size_t i = 0; i < impls.size(); ++i) {
        std::fprintf(stderr, "  %-6s --impl %-12s %s%s\n", op, impls[i].name, impls[i].summary,
                     i == 0 ? " (default)" : "");
    }
}

void usage()
{
    std::fprintf(stderr,
        "usage:\n"
        "  rfgpu spectrum  --in X.cf32 --out OUT.cf32 [--nfft N]\n"
        "  rfgpu fir       --in X.cf32 --taps H.f32 --out OUT.cf32 [--impl NAME] [--block B]\n"
        "  rfgpu xcorr     --a A.cf32 --b B.cf32 --out OUT.cf32 [--impl NAME]\n"
        "  rfgpu magnitude --in X.cf32 --out OUT.f32\n"
        "  rfgpu channelize --in X.cf32 --taps H.f32 --channels M --out OUT.cf32\n"
        "  rfgpu gcc_phat  --a A.cf32 --b B.cf32 --out OUT.cf32\n"
        "  rfgpu doa       --in CH0,CH1,... --spacing M [--speed M/S] [--rate HZ]\n"
        "  rfgpu stream    --in X.cf32 --taps H.f32 --out OUT.cf32 [--impl NAME] [--chunk N]\n"
        "                  [--depth D] [--memory copy|registered|zero-copy] [--rate HZ]\n"
        "  rfgpu info\n"
        "\n"
        "Inputs may be SigMF recordings (NAME.sigmf-meta / .sigmf-data): cf32_le, ci16_le, ci8.\n"
        "benchmark options (spectrum, magnitude, fir, xcorr, channelize, gcc_phat):\n"
        "  --bench N      time N runs with device events and print JSON\n"
        "  --warmup W     untimed runs first (default 3, needs --bench)\n"
        "fir options:\n"
        "  --block B      threads per workgroup, a multiple of 32 up to 1024 (default 256)\n"
        "\n"
        "implementations:\n");
    list_impls("fir", fir_impls());
    list_impls("xcorr", xcorr_impls());
}

class Args {
public:
    Args(int argc, char** argv, const std::set<std::string>& allowed)
    {
        for (int i = 2; i < argc; i += 2) {
            const std::string flag = argv[i];
            if (allowed.count(flag) == 0) {
                throw UsageError("unknown option '" + flag + "' for " + argv[1]);
            }
            if (i + 1 >= argc) {
                throw UsageError("option " + flag + " needs a value");
            }
            if (!values_.emplace(flag, argv[i + 1]).second) {
                throw UsageError("option " + flag + " given twice");
            }
        }
    }
    bool has(const std::string& f) const { return values_.count(f) != 0; }
    const std::string& required(const std::string& f) const
    {
        const auto it = values_.find(f);
        if (it == values_.end()) {
            throw UsageError("missing required option " + f);
        }
        return it->second;
    }
    std::string optional(const std::string& f, const std::string& fallback) const
    {
        const auto it = values_.find(f);
        return it == values_.end() ? fallback : it->second;
    }
    int integer(const std::string& f, int fallback, int min) const
    {
        const auto it = values_.find(f);
        if (it == values_.end()) {
            return fallback;
        }
        std::size_t used = 0;
        int v = 0;
        try {
            v = std::stoi(it->second, &used);
        } catch (const std::exception&) {
            used = 0;
        }
        if (used == 0 || used != it->second.size() || v < min) {
            throw UsageError("option " + f + " needs an integer >= " + std::to_string(min)
                             + ", got '" + it->second + "'");
        }
        return v;
    }

private:
    std::map<std::string, std::string> values_;
};

template <class Fn>
const Impl<Fn>& pick(const std::vector<Impl<Fn>>& impls, const Args& args, const char* op)
{
    const std::string name = args.optional("--impl", impls.front().name);
    for (const auto& i : impls) {
        if (name == i.name) {
            return i;
        }
    }
    throw UsageError("unknown --impl '" + name + "' for " + op + " (available: "
                     + names_of(impls, ", ") + ")");
}

BenchConfig bench_config(const Args& args)
{
    BenchConfig c;
    if (args.has("--bench")) {
        c.runs = args.integer("--bench", 1, 1);
        c.warmup = args.integer("--warmup", 3, 0);
        if (c.runs < 10) {
            std::fprintf(stderr, "warning: --bench %d is below 10 runs\n", c.runs);
        }
    } else if (args.has("--warmup")) {
        throw UsageError("--warmup needs --bench");
    }
    return c;
}

void require_int_range(std::size_t count, const std::string& what)
{
    if (count > static_cast<std::size_t>(INT_MAX)) {
        throw std::runtime_error(what + " has " + std::to_string(count)
                                 + " elements, more than the supported " + std::to_string(INT_MAX));
    }
}

std::string gpu_name()
{
    char buf[256];
    RF_CHECK(rf_device_name(buf, sizeof buf));
    return buf;
}

void report(const Args& args, const char* op, const char* impl, std::size_t num_in,
            std::size_t num_out, const std::string& path, const Bench& bench)
{
    if (!args.has("--bench")) {
        std::printf("%s[%s]: %zu samples in, %zu out -> %s\n", op, impl, num_in, num_out,
                    path.c_str());
        return;
    }
    const BenchResult& r = bench.result();
    std::printf("{\"op\": \"%s\", \"impl\": \"%s\", \"gpu\": \"%s\", \"num_in\": %zu, "
                "\"num_out\": %zu, \"runs\": %d, \"warmup\": %d, "
                "\"h2d_ms\": {\"mean\": %.6f, \"std\": %.6f}, "
                "\"kernel_ms\": {\"mean\": %.6f, \"std\": %.6f}, "
                "\"d2h_ms\": {\"mean\": %.6f, \"std\": %.6f}}\n",
                op, impl, gpu_name().c_str(), num_in, num_out, bench.config().runs,
                bench.config().warmup, r.h2d.mean_ms, r.h2d.std_ms, r.kernel.mean_ms,
                r.kernel.std_ms, r.d2h.mean_ms, r.d2h.std_ms);
}

template <class T>
void finish(const Args& args, const char* op, const char* impl, std::size_t num_in,
            const std::vector<T>& out, const Bench& bench)
{
    const std::string& path = args.required("--out");
    if constexpr (sizeof(T) == sizeof(cf32)) {
        write_cf32(path, out);
    } else {
        write_f32(path, out);
    }
    report(args, op, impl, num_in, out.size(), path, bench);
}

int run_spectrum(int argc, char** argv)
{
    const Args args(argc, argv, {"--in", "--out", "--nfft", "--bench", "--warmup"});
    args.required("--out");
    const int nfft = args.integer("--nfft", 1024, 1);
    Bench bench(bench_config(args));
    const std::vector<cf32> x = read_cf32(args.required("--in"));
    require_int_range(x.size() + static_cast<std::size_t>(nfft), "padded spectrum input");
    finish(args, "spectrum", "fft", x.size(), spectrum_fft(x, nfft, bench), bench);
    return 0;
}

int run_magnitude(int argc, char** argv)
{
    const Args args(argc, argv, {"--in", "--out", "--bench", "--warmup"});
    args.required("--out");
    Bench bench(bench_config(args));
    const std::vector<cf32> x = read_cf32(args.required("--in"));
    require_int_range(x.size(), "magnitude input");
    const std::vector<float> mag = x.empty() ? std::vector<float>() : magnitude_hip(x, bench);
    finish(args, "magnitude", "elementwise", x.size(), mag, bench);
    return 0;
}

int run_fir(int argc, char** argv)
{
    const Args args(argc, argv,
                    {"--in", "--taps", "--out", "--impl", "--block", "--bench", "--warmup"});
    args.required("--out");
    const Impl<FirFn>& impl = pick(fir_impls(), args, "fir");
    const int block = args.integer("--block", kDefaultFirBlock, 32);
    if (block % 32 != 0 || block > 1024) {
        throw UsageError("--block needs a multiple of 32 up to 1024, got " + std::to_string(block));
    }
    Bench bench(bench_config(args));
    const std::vector<cf32> x = read_cf32(args.required("--in"));
    const std::vector<float> taps = read_f32(args.required("--taps"));
    if (taps.empty()) {
        throw std::runtime_error(args.required("--taps") + ": no taps");
    }
    require_int_range(x.size(), "fir input");
    require_int_range(taps.size(), "tap array");
    const std::vector<cf32> y = x.empty() ? std::vector<cf32>() : impl.fn(x, taps, block, bench);
    finish(args, "fir", impl.name, x.size(), y, bench);
    return 0;
}

int run_xcorr(int argc, char** argv)
{
    const Args args(argc, argv, {"--a", "--b", "--out", "--impl", "--bench", "--warmup"});
    args.required("--out");
    const Impl<XcorrFn>& impl = pick(xcorr_impls(), args, "xcorr");
    Bench bench(bench_config(args));
    const std::vector<cf32> a = read_cf32(args.required("--a"));
    const std::vector<cf32> b = read_cf32(args.required("--b"));
    if (a.empty() || b.empty()) {
        throw std::runtime_error("xcorr needs at least one sample in each input");
    }
    require_int_range(a.size() + b.size(), "xcorr output");
    finish(args, "xcorr", impl.name, a.size() + b.size(), impl.fn(a, b, bench), bench);
    return 0;
}

int run_channelize(int argc, char** argv)
{
    const Args args(argc, argv, {"--in", "--taps", "--channels", "--out", "--bench", "--warmup"});
    args.required("--out");
    const int channels = args.integer("--channels", 0, 1);
    if (!args.has("--channels")) {
        throw UsageError("missing required option --channels");
    }
    Bench bench(bench_config(args));
    const std::vector<cf32> x = read_cf32(args.required("--in"));
    const std::vector<float> taps = read_f32(args.required("--taps"));
    if (taps.empty()) {
        throw std::runtime_error(args.required("--taps") + ": no taps");
    }
    require_int_range(x.size() + static_cast<std::size_t>(channels), "padded channelizer input");
    require_int_range(taps.size(), "tap array");
    finish(args, "channelize", "polyphase", x.size(), channelize_hip(x, taps, channels, bench),
           bench);
    return 0;
}

int run_gcc_phat(int argc, char** argv)
{
    const Args args(argc, argv, {"--a", "--b", "--out", "--bench", "--warmup"});
    args.required("--out");
    Bench bench(bench_config(args));
    const std::vector<cf32> a = read_cf32(args.required("--a"));
    const std::vector<cf32> b = read_cf32(args.required("--b"));
    if (a.empty() || b.empty()) {
        throw std::runtime_error("gcc_phat needs at least one sample in each input");
    }
    require_int_range(a.size() + b.size(), "gcc_phat output");
    finish(args, "gcc_phat", "fft", a.size() + b.size(), gcc_phat_hip(a, b, bench), bench);
    return 0;
}

// Direction of arrival for a uniform linear array: GCC-PHAT of every channel against channel
// 0 with the peak found on the device, then the least-squares angle (src/main.cpp:352-416).
int run_doa(int argc, char** argv)
{
    const Args args(argc, argv, {"--in", "--spacing", "--speed", "--rate"});
    std::vector<std::string> paths;
    const std::string list = args.required("--in");
    for (std::size_t start = 0; start <= list.size();) {
        const std::size_t comma = list.find(',', start);
        const std::size_t end = comma == std::string::npos ? list.size() : comma;
        paths.push_back(list.substr(start, end - start));
        start = end + 1;
    }
    if (paths.size() < 2) {
        throw UsageError("--in needs at least two comma-separated channel files");
    }
    const double spacing = std::stod(args.required("--spacing"));
    const double speed = std::stod(args.optional("--speed", "299792458"));
    double rate = sample_rate(paths[0]);
    if (args.has("--rate")) {
        rate = std::stod(args.required("--rate"));
    }
    if (!(spacing > 0.0) || !(speed > 0.0) || !(rate > 0.0)) {
        throw UsageError("--spacing, --speed, and the sample rate (--rate, or the first "
                         "channel's metadata) must be positive");
    }
    const std::vector<cf32> ref = read_cf32(paths[0]);
    std::vector<int> lags;
    double sum_it = 0.0;
    double sum_ii = 0.0;
    for (std::size_t i = 1; i < paths.size(); ++i) {
        const std::vector<cf32> ch = read_cf32(paths[i]);
        if (ref.empty() || ch.empty()) {
            throw std::runtime_error("doa: every channel needs at least one sample");
        }
        const int lag = tdoa_hip(ref, ch);
        lags.push_back(lag);
        sum_it += static_cast<double>(i) * lag / rate;
        sum_ii += static_cast<double>(i * i);
    }
    const double sine = std::clamp(speed * (sum_it / sum_ii) / spacing, -1.0, 1.0);
    const double angle = std::asin(sine) * 180.0 / 3.14159265358979323846;
    std::string text;
    for (int lag : lags) {
        text += (text.empty() ? "" : ", ") + std::to_string(lag);
    }
    std::printf("{\"op\": \"doa\", \"channels\": %zu, \"sample_rate\": %.1f, \"spacing_m\": %.6f, "
                "\"speed_m_s\": %.1f, \"tdoa_samples\": [%s], \"angle_deg\": %.9f}\n",
                paths.size(), rate, spacing, speed, text.c_str(), angle);
    return 0;
}

int run_stream(int argc, char** argv)
{
    const Args args(argc, argv, {"--in", "--taps", "--out", "--impl", "--chunk", "--depth",
                                 "--memory", "--rate", "--bench", "--warmup"});
    args.required("--out");
    StreamConfig config;
    const std::string impl = args.optional("--impl", "tiled");
    if (!parse_fir_impl(impl.c_str(), config.impl)) {
        throw UsageError("unknown --impl '" + impl + "' for stream (available: "
                         + names_of(fir_impls(), ", ") + ")");
    }
    const std::string memory = args.optional("--memory", "copy");
    if (!parse_stream_memory(memory, config.memory)) {
        throw UsageError("unknown --memory '" + memory + "' (available: copy, registered, zero-copy)");
    }
    config.chunk = args.integer("--chunk", config.chunk, 1);
    config.depth = args.integer("--depth", config.depth, 1);
    const BenchConfig bench = bench_config(args);
    const std::vector<cf32> x = read_cf32(args.required("--in"));
    const std::vector<float> taps = read_f32(args.required("--taps"));
    if (taps.empty()) {
        throw std::runtime_error(args.required("--taps") + ": no taps");
    }
    require_int_range(x.size(), "stream input");
    require_int_range(static_cast<std::size_t>(config.chunk) + taps.size(), "chunk with history");
    double rate = sample_rate(args.required("--in"));
    if (args.has("--rate")) {
        rate = std::stod(args.required("--rate"));
    }
    std::vector<double> times;
    std::vector<cf32> y;
    StreamTiming timing;
    for (int i = 0; i < bench.warmup + bench.runs; ++i) {
        y = fir_stream(x, taps, config, timing);
        if (i >= bench.warmup) {
            times.push_back(timing.total_ms);
        }
    }
    write_cf32(args.required("--out"), y);
    if (!args.has("--bench")) {
        std::printf("stream[%s]: %zu samples in %d chunks of %d, depth %d -> %s\n", impl.c_str(),
                    x.size(), timing.chunks, config.chunk, config.depth,
                    args.required("--out").c_str());
        return 0;
    }
    double mean = 0.0;
    for (double t : times) {
        mean += t / static_cast<double>(times.size());
    }
    double var = 0.0;
    for (double t : times) {
        var += (t - mean) * (t - mean);
    }
    const double sd = times.size() > 1 ? std::sqrt(var / static_cast<double>(times.size() - 1)) : 0.0;
    const double msps = static_cast<double>(x.size()) / (mean * 1e3);
    std::printf("{\"op\": \"stream\", \"impl\": \"%s\", \"memory\": \"%s\", \"gpu\": \"%s\", "
                "\"num_in\": %zu, \"chunk\": %d, \"depth\": %d, \"chunks\": %d, \"runs\": %d, "
                "\"warmup\": %d, \"total_ms\": {\"mean\": %.6f, \"std\": %.6f}, \"msps\": %.3f, "
                "\"sample_rate\": %.1f, \"realtime_factor\": %.3f}\n",
                impl.c_str(), memory.c_str(), gpu_name().c_str(), x.size(), config.chunk,
                config.depth, timing.chunks, bench.runs, bench.warmup, mean, sd, msps, rate,
                rate > 0 ? msps * 1e6 / rate : 0.0);
    return 0;
}

int run_info(int argc, char** argv)
{
    const Args args(argc, argv, {});
    size_t memory = 0;
    RF_CHECK(rf_device_memory(&memory));
    std::printf("{\"gpu\": \"%s\", \"memory_mb\": %zu, \"abi\": %d, "
                "\"impls\": {\"spectrum\": [\"fft\"], \"magnitude\": [\"elementwise\"], "
                "\"fir\": [\"%s\"], \"xcorr\": [\"%s\"]}}\n",
                gpu_name().c_str(), memory / (1024 * 1024), rf_abi_version(),
                names_of(fir_impls(), "\", \"").c_str(), names_of(xcorr_impls(), "\", \"").c_str());
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        const std::string op = argc > 1 ? argv[1] : "";
        if (op == "spectrum") return run_spectrum(argc, argv);
        if (op == "magnitude") return run_magnitude(argc, argv);
        if (op == "fir") return run_fir(argc, argv);
        if (op == "xcorr") return run_xcorr(argc, argv);
        if (op == "stream") return run_stream(argc, argv);
        if (op == "channelize") return run_channelize(argc, argv);
        if (op == "gcc_phat") return run_gcc_phat(argc, argv);
        if (op == "doa") return run_doa(argc, argv);
        if (op == "info") return run_info(argc, argv);
        if (op == "--help" || op == "-h" || op == "help") {
            usage();
            return 0;
        }
        throw UsageError(op.empty() ? "missing subcommand" : "unknown subcommand '" + op + "'");
    } catch (const UsageError& e) {
        std::fprintf(stderr, "rfgpu: %s\n\n", e.what());
        usage();
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "rfgpu: %s\n", e.what());
        return 1;
    }
}
