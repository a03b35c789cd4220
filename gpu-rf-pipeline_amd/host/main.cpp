// rfgpu: command-line front end of the MI355X hot path. Its interface is the reference's
// (src/main.cpp): the subcommands, their flags, the JSON reports and the exit codes 0 (done),
// 1 (runtime error: files, device) and 2 (usage error).
//
// Every subcommand is a row of kCommands: its flags (name, type, lower bound), a synopsis for
// the usage text, and a handler that receives the parsed, type-checked options. Reports are
// built with a small JSON writer.

#include "bench.hpp"
#include "check.hpp"
#include "io.hpp"
#include "ops.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace {

struct UsageError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// --- JSON reports --------------------------------------------------------------------------

// A JSON object built field by field, in insertion order.
class Json {
public:
    Json& add(const std::string& key, const std::string& text)
    {
        return raw(key, quote(text));
    }
    Json& add(const std::string& key, const char* text) { return add(key, std::string(text)); }
    Json& add(const std::string& key, long long v) { return raw(key, std::to_string(v)); }
    Json& add(const std::string& key, int v) { return add(key, static_cast<long long>(v)); }
    Json& add(const std::string& key, std::size_t v)
    {
        return add(key, static_cast<long long>(v));
    }
    Json& add(const std::string& key, double v, int decimals)
    {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*f", decimals, std::isfinite(v) ? v : 0.0);
        return raw(key, buf);
    }
    Json& add(const std::string& key, const Json& object) { return raw(key, object.str()); }
    Json& add(const std::string& key, const std::vector<std::string>& list)
    {
        std::string out;
        for (const std::string& s : list) {
            out += (out.empty() ? "" : ", ") + quote(s);
        }
        return raw(key, "[" + out + "]");
    }
    Json& add(const std::string& key, const std::vector<int>& list)
    {
        std::string out;
        for (int v : list) {
            out += (out.empty() ? "" : ", ") + std::to_string(v);
        }
        return raw(key, "[" + out + "]");
    }
    std::string str() const { return "{" + body_ + "}"; }
    void print() const { std::printf("%s\n", str().c_str()); }

private:
    static std::string quote(const std::string& s)
    {
        std::string out = "\"";
        for (char c : s) {
            if (c == '"' || c == '\\') {
                out += '\\';
            }
            out += c;
        }
        return out + "\"";
    }
    Json& raw(const std::string& key, const std::string& value)
    {
        body_ += (body_.empty() ? "" : ", ") + quote(key) + ": " + value;
        return *this;
    }
    std::string body_;
};

Json phase(const PhaseTime& t)
{
    return Json().add("mean", t.mean_ms, 6).add("std", t.std_ms, 6);
}

// --- flags ---------------------------------------------------------------------------------

enum class Type { Text, Count, Real };

struct Flag {
    const char* name;
    Type type;
    int min;  // Count: smallest accepted value
};

// The flags of one invocation, checked against the command's table as they are read.
class Options {
public:
    Options(const std::string& command, const std::vector<Flag>& table, int argc, char** argv)
    {
        for (int i = 2; i < argc; i += 2) {
            const std::string name = argv[i];
            const Flag* flag = nullptr;
            for (const Flag& f : table) {
                if (name == f.name) {
                    flag = &f;
                }
            }
            if (flag == nullptr) {
                throw UsageError("unknown option '" + name + "' for " + command);
            }
            if (i + 1 >= argc) {
                throw UsageError("option " + name + " needs a value");
            }
            const std::string value = argv[i + 1];
            if (flag->type == Type::Count) {
                check_count(*flag, value);
            } else if (flag->type == Type::Real) {
                check_real(*flag, value);
            }
            if (!values_.emplace(name, value).second) {
                throw UsageError("option " + name + " given twice");
            }
        }
    }

    bool has(const std::string& name) const { return values_.count(name) != 0; }

    const std::string& need(const std::string& name) const
    {
        const auto it = values_.find(name);
        if (it == values_.end()) {
            throw UsageError("missing required option " + name);
        }
        return it->second;
    }
    std::string text(const std::string& name, const std::string& fallback) const
    {
        return has(name) ? need(name) : fallback;
    }
    int count(const std::string& name, int fallback) const
    {
        return has(name) ? std::atoi(need(name).c_str()) : fallback;
    }
    double real(const std::string& name, double fallback) const
    {
        return has(name) ? std::strtod(need(name).c_str(), nullptr) : fallback;
    }

private:
    static void check_count(const Flag& f, const std::string& v)
    {
        char* end = nullptr;
        errno = 0;
        const long n = v.empty() ? 0 : std::strtol(v.c_str(), &end, 10);
        if (v.empty() || *end != '\0' || errno != 0 || n < f.min || n > INT_MAX) {
            throw UsageError(std::string("option ") + f.name + " needs an integer >= "
                             + std::to_string(f.min) + ", got '" + v + "'");
        }
    }
    static void check_real(const Flag& f, const std::string& v)
    {
        char* end = nullptr;
        (void)std::strtod(v.c_str(), &end);
        if (v.empty() || *end != '\0') {
            throw UsageError(std::string("option ") + f.name + " needs a number, got '" + v
                             + "'");
        }
    }
    std::map<std::string, std::string> values_;
};

// --- shared pieces of the handlers --------------------------------------------------------

const std::vector<Flag> kBenchFlags = {{"--bench", Type::Count, 1}, {"--warmup", Type::Count, 0}};

BenchConfig bench_config(const Options& o)
{
    if (o.has("--warmup") && !o.has("--bench")) {
        throw UsageError("--warmup needs --bench");
    }
    BenchConfig c;
    if (o.has("--bench")) {
        c.runs = o.count("--bench", 1);
        c.warmup = o.count("--warmup", 3);
        if (c.runs < 10) {
            std::fprintf(stderr, "warning: --bench %d is below 10 runs\n", c.runs);
        }
    }
    return c;
}

// Sizes go to the C ABI as int.
void fits_int(std::size_t count, const std::string& what)
{
    if (count > static_cast<std::size_t>(INT_MAX)) {
        throw std::runtime_error(what + " has " + std::to_string(count)
                                 + " elements, more than the supported "
                                 + std::to_string(INT_MAX));
    }
}

std::vector<float> load_taps(const Options& o)
{
    std::vector<float> taps = read_f32(o.need("--taps"));
    if (taps.empty()) {
        throw std::runtime_error(o.need("--taps") + ": no taps");
    }
    fits_int(taps.size(), "tap array");
    return taps;
}

std::string gpu_name()
{
    char name[256] = {};
    RF_CHECK(rf_device_name(name, sizeof name));
    return name;
}

template <class Fn>
std::vector<std::string> impl_names(const std::vector<Impl<Fn>>& impls)
{
    std::vector<std::string> names;
    for (const auto& i : impls) {
        names.push_back(i.name);
    }
    return names;
}

std::string joined(const std::vector<std::string>& names)
{
    std::string out;
    for (const std::string& n : names) {
        out += (out.empty() ? "" : ", ") + n;
    }
    return out;
}

template <class Fn>
const Impl<Fn>& choose(const std::vector<Impl<Fn>>& impls, const Options& o, const char* op)
{
    const std::string want = o.text("--impl", impls.front().name);
    const auto it = std::find_if(impls.begin(), impls.end(),
                                 [&](const Impl<Fn>& i) { return want == i.name; });
    if (it == impls.end()) {
        throw UsageError("unknown --impl '" + want + "' for " + op + " (available: "
                         + joined(impl_names(impls)) + ")");
    }
    return *it;
}

// Writes the result and reports: a one-line summary, or the bench JSON with --bench.
template <class T>
int finish(const Options& o, const char* op, const char* impl, std::size_t num_in,
           const std::vector<T>& out, const Bench& bench)
{
    const std::string& path = o.need("--out");
    if constexpr (std::is_same_v<T, cf32>) {
        write_cf32(path, out);
    } else {
        write_f32(path, out);
    }
    if (!o.has("--bench")) {
        std::printf("%s[%s]: %zu samples in, %zu out -> %s\n", op, impl, num_in, out.size(),
                    path.c_str());
        return 0;
    }
    const BenchResult& r = bench.result();
    Json()
        .add("op", op)
        .add("impl", impl)
        .add("gpu", gpu_name())
        .add("num_in", num_in)
        .add("num_out", out.size())
        .add("runs", bench.config().runs)
        .add("warmup", bench.config().warmup)
        .add("h2d_ms", phase(r.h2d))
        .add("kernel_ms", phase(r.kernel))
        .add("d2h_ms", phase(r.d2h))
        .print();
    return 0;
}

// --- subcommands ----------------------------------------------------------------------------

int cmd_spectrum(const Options& o)
{
    o.need("--out");
    const int nfft = o.count("--nfft", 1024);
    Bench bench(bench_config(o));
    const std::vector<cf32> x = read_cf32(o.need("--in"));
    fits_int(x.size() + static_cast<std::size_t>(nfft), "padded spectrum input");
    return finish(o, "spectrum", "fft", x.size(), spectrum_fft(x, nfft, bench), bench);
}

int cmd_magnitude(const Options& o)
{
    o.need("--out");
    Bench bench(bench_config(o));
    const std::vector<cf32> x = read_cf32(o.need("--in"));
    fits_int(x.size(), "magnitude input");
    const std::vector<float> mag = x.empty() ? std::vector<float>{} : magnitude_hip(x, bench);
    return finish(o, "magnitude", "elementwise", x.size(), mag, bench);
}

int cmd_fir(const Options& o)
{
    o.need("--out");
    const Impl<FirFn>& impl = choose(fir_impls(), o, "fir");
    const int block = o.count("--block", kDefaultFirBlock);
    if (block % 32 != 0 || block > 1024) {
        throw UsageError("--block needs a multiple of 32 up to 1024, got "
                         + std::to_string(block));
    }
    Bench bench(bench_config(o));
    const std::vector<cf32> x = read_cf32(o.need("--in"));
    const std::vector<float> taps = load_taps(o);
    fits_int(x.size(), "fir input");
    const std::vector<cf32> y = x.empty() ? std::vector<cf32>{} : impl.fn(x, taps, block, bench);
    return finish(o, "fir", impl.name, x.size(), y, bench);
}

// xcorr and gcc_phat read two captures; both need a sample in each.
std::pair<std::vector<cf32>, std::vector<cf32>> read_pair(const Options& o, const char* op)
{
    std::pair<std::vector<cf32>, std::vector<cf32>> ab{read_cf32(o.need("--a")),
                                                       read_cf32(o.need("--b"))};
    if (ab.first.empty() || ab.second.empty()) {
        throw std::runtime_error(std::string(op) + " needs at least one sample in each input");
    }
    fits_int(ab.first.size() + ab.second.size(), std::string(op) + " output");
    return ab;
}

int cmd_xcorr(const Options& o)
{
    o.need("--out");
    const Impl<XcorrFn>& impl = choose(xcorr_impls(), o, "xcorr");
    Bench bench(bench_config(o));
    const auto [a, b] = read_pair(o, "xcorr");
    return finish(o, "xcorr", impl.name, a.size() + b.size(), impl.fn(a, b, bench), bench);
}

int cmd_gcc_phat(const Options& o)
{
    o.need("--out");
    Bench bench(bench_config(o));
    const auto [a, b] = read_pair(o, "gcc_phat");
    return finish(o, "gcc_phat", "fft", a.size() + b.size(), gcc_phat_hip(a, b, bench), bench);
}

int cmd_channelize(const Options& o)
{
    o.need("--out");
    const int channels = std::atoi(o.need("--channels").c_str());
    Bench bench(bench_config(o));
    const std::vector<cf32> x = read_cf32(o.need("--in"));
    const std::vector<float> taps = load_taps(o);
    fits_int(x.size() + static_cast<std::size_t>(channels), "padded channelizer input");
    return finish(o, "channelize", "polyphase", x.size(),
                  channelize_hip(x, taps, channels, bench), bench);
}

std::vector<std::string> split_list(const std::string& list)
{
    std::vector<std::string> items;
    std::size_t from = 0;
    for (;;) {
        const std::size_t comma = list.find(',', from);
        items.push_back(list.substr(from, comma == std::string::npos ? std::string::npos
                                                                    : comma - from));
        if (comma == std::string::npos) {
            return items;
        }
        from = comma + 1;
    }
}

// Direction of arrival for a uniform linear array (interface: src/main.cpp:352-416): the lag of
// every channel against channel 0 from one batched GCC-PHAT with the peak search on the device,
// then the least-squares slope of lag time against element index, sin(angle) = c slope / d.
int cmd_doa(const Options& o)
{
    const std::vector<std::string> paths = split_list(o.need("--in"));
    if (paths.size() < 2) {
        throw UsageError("--in needs at least two comma-separated channel files");
    }
    o.need("--spacing");
    const double spacing = o.real("--spacing", 0.0);
    const double speed = o.real("--speed", 299792458.0);
    const double rate = o.has("--rate") ? o.real("--rate", 0.0) : sample_rate(paths[0]);
    if (!(spacing > 0.0) || !(speed > 0.0) || !(rate > 0.0)) {
        throw UsageError("--spacing, --speed, and the sample rate (--rate, or the first "
                         "channel's metadata) must be positive");
    }
    std::vector<std::vector<cf32>> chans;
    for (const std::string& p : paths) {
        chans.push_back(read_cf32(p));
        if (chans.back().empty()) {
            throw std::runtime_error("doa: every channel needs at least one sample");
        }
        fits_int(chans.front().size() + chans.back().size(), "doa pair output");
    }
    const std::vector<int> lags = tdoa_batch_hip(chans);
    double num = 0.0;
    double den = 0.0;
    for (std::size_t i = 1; i < chans.size(); ++i) {
        num += static_cast<double>(i) * (lags[i - 1] / rate);
        den += static_cast<double>(i) * static_cast<double>(i);
    }
    const double s = std::max(-1.0, std::min(1.0, speed * (num / den) / spacing));
    const double degrees = std::asin(s) * (180.0 / 3.14159265358979323846);
    Json()
        .add("op", "doa")
        .add("channels", chans.size())
        .add("sample_rate", rate, 1)
        .add("spacing_m", spacing, 6)
        .add("speed_m_s", speed, 1)
        .add("tdoa_samples", lags)
        .add("angle_deg", degrees, 9)
        .print();
    return 0;
}

int cmd_stream(const Options& o)
{
    o.need("--out");
    StreamConfig config;
    const std::string impl = o.text("--impl", "tiled");
    if (!parse_fir_impl(impl.c_str(), config.impl)) {
        throw UsageError("unknown --impl '" + impl + "' for stream (available: "
                         + joined(impl_names(fir_impls())) + ")");
    }
    const std::string memory = o.text("--memory", "copy");
    if (!parse_stream_memory(memory, config.memory)) {
        throw UsageError("unknown --memory '" + memory
                         + "' (available: copy, registered, zero-copy)");
    }
    config.chunk = o.count("--chunk", config.chunk);
    config.depth = o.count("--depth", config.depth);
    const BenchConfig bench = bench_config(o);
    const std::vector<cf32> x = read_cf32(o.need("--in"));
    const std::vector<float> taps = load_taps(o);
    fits_int(x.size(), "stream input");
    fits_int(static_cast<std::size_t>(config.chunk) + taps.size(), "chunk with history");
    const double rate = o.has("--rate") ? o.real("--rate", 0.0) : sample_rate(o.need("--in"));

    std::vector<cf32> y;
    StreamTiming timing;
    double mean = 0.0;
    double m2 = 0.0;
    for (int i = -bench.warmup; i < bench.runs; ++i) {
        y = fir_stream(x, taps, config, timing);
        if (i >= 0) {
            const double d = timing.total_ms - mean;
            mean += d / (i + 1);
            m2 += d * (timing.total_ms - mean);
        }
    }
    write_cf32(o.need("--out"), y);
    if (!o.has("--bench")) {
        std::printf("stream[%s]: %zu samples in %d chunks of %d, depth %d -> %s\n", impl.c_str(),
                    x.size(), timing.chunks, config.chunk, config.depth, o.need("--out").c_str());
        return 0;
    }
    const double sd = bench.runs > 1 ? std::sqrt(m2 / (bench.runs - 1)) : 0.0;
    const double msps = static_cast<double>(x.size()) / (mean * 1e3);
    Json()
        .add("op", "stream")
        .add("impl", impl)
        .add("memory", memory)
        .add("gpu", gpu_name())
        .add("num_in", x.size())
        .add("chunk", config.chunk)
        .add("depth", config.depth)
        .add("chunks", timing.chunks)
        .add("runs", bench.runs)
        .add("warmup", bench.warmup)
        .add("total_ms", Json().add("mean", mean, 6).add("std", sd, 6))
        .add("msps", msps, 3)
        .add("sample_rate", rate, 1)
        .add("realtime_factor", rate > 0.0 ? msps * 1e6 / rate : 0.0, 3)
        .print();
    return 0;
}

int cmd_info(const Options&)
{
    std::size_t memory = 0;
    RF_CHECK(rf_device_memory(&memory));
    Json()
        .add("gpu", gpu_name())
        .add("memory_mb", memory / (1024 * 1024))
        .add("abi", rf_abi_version())
        .add("impls", Json()
                          .add("spectrum", std::vector<std::string>{"fft"})
                          .add("magnitude", std::vector<std::string>{"elementwise"})
                          .add("fir", impl_names(fir_impls()))
                          .add("xcorr", impl_names(xcorr_impls())))
        .print();
    return 0;
}

struct Command {
    const char* name;
    const char* synopsis;
    std::vector<Flag> flags;
    bool benchable;
    int (*run)(const Options&);
};

const std::vector<Command>& commands()
{
    static const std::vector<Command> kCommands = {
        {"spectrum", "--in X.cf32 --out OUT.cf32 [--nfft N]",
         {{"--in", Type::Text, 0}, {"--out", Type::Text, 0}, {"--nfft", Type::Count, 1}},
         true, cmd_spectrum},
        {"fir", "--in X.cf32 --taps H.f32 --out OUT.cf32 [--impl NAME] [--block B]",
         {{"--in", Type::Text, 0}, {"--taps", Type::Text, 0}, {"--out", Type::Text, 0},
          {"--impl", Type::Text, 0}, {"--block", Type::Count, 32}},
         true, cmd_fir},
        {"xcorr", "--a A.cf32 --b B.cf32 --out OUT.cf32 [--impl NAME]",
         {{"--a", Type::Text, 0}, {"--b", Type::Text, 0}, {"--out", Type::Text, 0},
          {"--impl", Type::Text, 0}},
         true, cmd_xcorr},
        {"magnitude", "--in X.cf32 --out OUT.f32",
         {{"--in", Type::Text, 0}, {"--out", Type::Text, 0}}, true, cmd_magnitude},
        {"channelize", "--in X.cf32 --taps H.f32 --channels M --out OUT.cf32",
         {{"--in", Type::Text, 0}, {"--taps", Type::Text, 0}, {"--channels", Type::Count, 1},
          {"--out", Type::Text, 0}},
         true, cmd_channelize},
        {"gcc_phat", "--a A.cf32 --b B.cf32 --out OUT.cf32",
         {{"--a", Type::Text, 0}, {"--b", Type::Text, 0}, {"--out", Type::Text, 0}}, true,
         cmd_gcc_phat},
        {"doa", "--in CH0,CH1,... --spacing M [--speed M/S] [--rate HZ]",
         {{"--in", Type::Text, 0}, {"--spacing", Type::Real, 0}, {"--speed", Type::Real, 0},
          {"--rate", Type::Real, 0}},
         false, cmd_doa},
        {"stream",
         "--in X.cf32 --taps H.f32 --out OUT.cf32 [--impl NAME] [--chunk N] [--depth D]\n"
         "                  [--memory copy|registered|zero-copy] [--rate HZ]",
         {{"--in", Type::Text, 0}, {"--taps", Type::Text, 0}, {"--out", Type::Text, 0},
          {"--impl", Type::Text, 0}, {"--chunk", Type::Count, 1}, {"--depth", Type::Count, 1},
          {"--memory", Type::Text, 0}, {"--rate", Type::Real, 0}},
         true, cmd_stream},
        {"info", "", {}, false, cmd_info},
    };
    return kCommands;
}

template <class Fn>
void print_impls(const char* op, const std::vector<Impl<Fn>>& impls)
{
    for (std::size_t i = 0; i < impls.size(); ++i) {
        std::fprintf(stderr, "  %-6s --impl %-12s %s%s\n", op, impls[i].name, impls[i].summary,
                     i == 0 ? " (default)" : "");
    }
}

void usage()
{
    std::fprintf(stderr, "usage:\n");
    for (const Command& c : commands()) {
        std::fprintf(stderr, "  rfgpu %-11s%s\n", c.name, c.synopsis);
    }
    std::fprintf(stderr,
                 "\nInputs may be SigMF recordings (NAME.sigmf-meta / .sigmf-data): cf32_le, "
                 "ci16_le, ci8.\n"
                 "--bench N       time N runs with device events and print JSON "
                 "(all but doa and info)\n"
                 "--warmup W      untimed runs first (default 3, needs --bench)\n"
                 "--block B       fir: threads per workgroup, a multiple of 32 up to 1024 "
                 "(default 256)\n\nimplementations:\n");
    print_impls("fir", fir_impls());
    print_impls("xcorr", xcorr_impls());
}

int dispatch(int argc, char** argv)
{
    const std::string name = argc > 1 ? argv[1] : "";
    if (name == "help" || name == "--help" || name == "-h") {
        usage();
        return 0;
    }
    for (const Command& c : commands()) {
        if (name == c.name) {
            std::vector<Flag> flags = c.flags;
            if (c.benchable) {
                flags.insert(flags.end(), kBenchFlags.begin(), kBenchFlags.end());
            }
            return c.run(Options(name, flags, argc, argv));
        }
    }
    throw UsageError(name.empty() ? "missing subcommand" : "unknown subcommand '" + name + "'");
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return dispatch(argc, argv);
    } catch (const UsageError& e) {
        std::fprintf(stderr, "rfgpu: %s\n\n", e.what());
        usage();
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "rfgpu: %s\n", e.what());
        return 1;
    }
}
