// The HIP backend of sigflowc (--emit=hip): a C++ program that runs one SigFlow pipeline on the
// GPU through host/device_ops.hpp and host/pipeline_runtime.hpp. The reference's CUDA backend
// (compiler/src/backend_cuda.cpp:36-374) sets the behaviour: captures read once, every value in
// the GPU buffer slot liveness analysis gave it, taps and FFT plans made before timing, a fusion
// group's FIRs as one batched call, emitted values copied back and written after the run.
// What is different here, by design for one MI355X:
//   - all work, copies included, goes on the runtime's one stream, so the compute phase can be
//     recorded once as a hipGraph (--graph) and replayed without per-call launch overhead;
//   - captures and results live in page-locked host memory: copies are asynchronous DMA;
//   - an FFT input's last-frame padding is cleared by a fill kernel, which a graph can hold;
//   - fft |> magnitude with nothing else reading the spectrum is one spectrum-magnitude
//     call (fuse_spectrum): the complex spectrum is never stored (up to 4096 points it never
//     leaves the kernel; longer transforms keep only the first pass in the plan's scratch).

#include "sigflow.hpp"

#include <set>

namespace sigflow {

const std::vector<std::string>& fir_impl_names()
{
    static const std::vector<std::string> names = {"naive", "tiled", "tiled_const", "register",
                                                   "fft_ola"};
    return names;
}

const std::vector<std::string>& xcorr_impl_names()
{
    static const std::vector<std::string> names = {"naive", "tiled", "fft"};
    return names;
}

namespace {

// A C++ string literal of `text`: quotes and backslashes escaped, control characters as
// three-digit octal escapes (a fixed length, so a following digit cannot extend them).
std::string quote(const std::string& text)
{
    std::string out = "\"";
    for (char c : text) {
        const unsigned char u = static_cast<unsigned char>(c);
        if (c == '"' || c == '\\') {
            out += '\\';
            out += c;
        } else if (u < 0x20 || u == 0x7f) {
            const char octal[] = {'\\', static_cast<char>('0' + (u >> 6)),
                                  static_cast<char>('0' + ((u >> 3) & 7)),
                                  static_cast<char>('0' + (u & 7)), '\0'};
            out += octal;
        } else {
            out += c;
        }
    }
    return out + "\"";
}

// Text of a node as a comment: no line breaks, whatever a path holds.
std::string comment(const Node& n)
{
    std::string s = node_text(n);
    for (char& c : s) {
        if (c == '\n' || c == '\r') {
            c = ' ';
        }
    }
    return s;
}

const char* fir_enum(const std::string& impl)
{
    if (impl == "tiled") return "FirImpl::Tiled";
    if (impl == "tiled_const") return "FirImpl::TiledConst";
    if (impl == "register") return "FirImpl::Register";
    if (impl == "fft_ola") return "FirImpl::FftOla";
    return "FirImpl::Naive";
}

class HipEmitter {
public:
    HipEmitter(const Graph& g, const Schedule& s, const Liveness& l, const HipOptions& opt,
               std::ostream& out)
        : g_(g), s_(s), l_(l), opt_(opt), out_(out)
    {
    }

    void run(const std::string& program)
    {
        for (int id : s_.order) {
            (node(id).op == Op::Capture ? captures_ : node(id).op == Op::Emit ? emits_ : values_)
                .push_back(id);
        }
        bool fused = false;
        for (int id : s_.order) {
            fused = fused || node(id).group >= 0 || node(id).spectrum_of >= 0;
        }
        const std::string kernels = "fir=" + opt_.kernels.fir + " xcorr=" + opt_.kernels.xcorr;
        out_ << "// Generated by sigflowc from " << program << "; do not edit.\n"
             << "// Build: gpu-rf-pipeline_amd/sigflow/build_program.sh THIS.cpp PROGRAM. Run it from "
                "the directory\n"
             << "// its paths are relative to: PROGRAM [--bench N] [--warmup W] [--graph]. Prints "
                "a JSON report.\n"
             << "#include \"device_ops.hpp\"\n"
             << "#include \"pipeline_runtime.hpp\"\n\n"
             << "#include <cstdio>\n#include <exception>\n\n"
             << "int main(int argc, char** argv)\n"
             << "try {\n"
             << "    sigflow::Runtime rt(" << quote(program) << ", " << quote(kernels) << ", "
             << (fused ? "true" : "false") << ", " << (opt_.reuse ? "true" : "false")
             << ", argc, argv);\n"
             << "    const rf_stream s = rt.stream();\n";
        setup();
        compute();
        finish();
        out_ << "} catch (const std::exception& e) {\n"
             << "    std::fprintf(stderr, \"error: %s\\n\", e.what());\n"
             << "    return 1;\n"
             << "}\n";
    }

private:
    const Node& node(int id) const { return g_.nodes[static_cast<std::size_t>(id)]; }
    static std::string n(int id) { return "n" + std::to_string(id); }
    static const char* elem(const Node& v) { return v.type.real ? "float" : "cf32"; }
    int slot_index(int id) const { return l_.slot_of[static_cast<std::size_t>(id)]; }
    std::string slot(int id) const
    {
        return "rt.slot<" + std::string(elem(node(id))) + ">(" + std::to_string(slot_index(id))
               + ")";
    }
    static std::string count(int id) { return "sigflow::samples(" + n(id) + ")"; }
    // Frame length of an FFT-reading node: the fft itself, or the magnitude it was fused into.
    static bool reads_frames(const Node& v)
    {
        return v.op == Op::Fft || (v.op == Op::Magnitude && v.spectrum_of >= 0);
    }

    void setup()
    {
        out_ << "\n    // Captures, read once into page-locked memory.\n";
        for (int id : captures_) {
            out_ << "    const sigflow::HostArray<cf32>& h" << id << " = rt.read_capture("
                 << quote(node(id).path) << ");  // " << comment(node(id)) << "\n";
        }

        out_ << "\n    // Samples in every value.\n";
        for (int id : s_.order) {
            const Node& v = node(id);
            switch (v.op) {
            case Op::Capture:
                out_ << "    const std::size_t " << n(id) << " = h" << id << ".size();\n";
                break;
            case Op::Fir:
                out_ << "    const std::size_t " << n(id) << " = " << n(v.inputs[0]) << ";\n";
                break;
            case Op::Magnitude:
                if (v.spectrum_of >= 0) {
                    out_ << "    const std::size_t " << n(id) << " = sigflow::padded_to_frames("
                         << n(v.inputs[0]) << ", " << v.type.n << ");\n";
                } else {
                    out_ << "    const std::size_t " << n(id) << " = " << n(v.inputs[0]) << ";\n";
                }
                break;
            case Op::Fft:
            case Op::Channelize:
                out_ << "    const std::size_t " << n(id) << " = sigflow::padded_to_frames("
                     << n(v.inputs[0]) << ", " << v.n << ");\n";
                break;
            case Op::Xcorr:
            case Op::GccPhat:
                out_ << "    const std::size_t " << n(id) << " = " << n(v.inputs[0]) << " + "
                     << n(v.inputs[1]) << " - 1;\n";
                break;
            case Op::Emit:
                break;
            }
        }

        out_ << "\n    // Buffer slots from liveness analysis, each as large as the largest value "
                "it holds, with the\n    // launches from the value's write to its last read "
                "(emitted values are read after the last).\n";
        for (const LiveRange& r : l_.ranges) {
            const Node& v = node(r.node);
            out_ << "    rt.need(" << r.slot << ", " << n(r.node) << " * sizeof(" << elem(v)
                 << "));  // %" << r.node << ", launches " << s_.launch[static_cast<std::size_t>(r.def)]
                 << ".." << s_.launch[static_cast<std::size_t>(r.last)] << "\n";
            for (int user : s_.order) {
                const Node& u = node(user);
                if (reads_frames(u) && u.inputs[0] == r.node) {
                    out_ << "    rt.need(" << r.slot << ", " << n(user) << " * sizeof(cf32));  // %"
                         << r.node << " padded for %" << user << "\n";
                }
            }
        }
        out_ << "    rt.allocate_slots();\n";

        bool first = true;
        const auto heading = [&] {
            if (first) {
                out_ << "\n    // Taps and plans.\n";
                first = false;
            }
        };
        std::set<int> groups_with_taps;  // a fusion group's batched call uses its first taps
        for (int id : s_.order) {
            const Node& v = node(id);
            const bool first_of_group = v.group < 0 || groups_with_taps.insert(v.group).second;
            if ((v.op == Op::Fir && first_of_group) || v.op == Op::Channelize) {
                heading();
                out_ << "    const sigflow::Taps t" << id << " = rt.taps(" << quote(v.taps) << ");\n";
            }
            if (v.op == Op::Channelize || reads_frames(v)) {
                heading();
                const long long frame = v.op == Op::Fft || v.op == Op::Channelize ? v.n : v.type.n;
                out_ << "    const rf_fft_plan* plan" << id << " = rt.fft_plan(" << frame << ", "
                     << n(id) << " / " << frame << ");\n";
            } else if (v.op == Op::GccPhat) {
                heading();
                out_ << "    const XcorrFft& gcc" << id << " = rt.xcorr_fft(" << n(v.inputs[0])
                     << ", " << n(v.inputs[1]) << ", true);\n";
            } else if (v.op == Op::Xcorr && opt_.kernels.xcorr == "fft") {
                heading();
                out_ << "    const XcorrFft& xcorr" << id << " = rt.xcorr_fft(" << n(v.inputs[0])
                     << ", " << n(v.inputs[1]) << ");\n";
            }
        }

        out_ << "\n    // Page-locked host copies of the emitted values.\n";
        for (int id : emits_) {
            const Node& value = node(node(id).inputs[0]);
            if (value.op != Op::Capture) {
                out_ << "    sigflow::HostArray<" << elem(value) << ">& out" << id << " = rt.host_"
                     << (value.type.real ? "f32" : "cf32") << "(" << n(value.id) << ");  // "
                     << comment(node(id)) << "\n";
            }
        }
    }

    void compute()
    {
        out_ << "\n    rt.run(\n        [&] {\n";
        for (int id : captures_) {
            out_ << "            rt.upload(" << slot_index(id) << ", h" << id << ");\n";
        }
        out_ << "        },\n        [&] {\n";
        for (int id : values_) {
            const Node& v = node(id);
            if (v.spectrum_of >= 0) {
                out_ << "            // " << comment(node(v.spectrum_of)) << "\n";
            }
            out_ << "            // " << comment(v) << "\n";
            switch (v.op) {
            case Op::Fir:
                if (v.group >= 0) {
                    fused_firs(v);
                } else {
                    out_ << "            fir_device(" << fir_enum(opt_.kernels.fir) << ", "
                         << slot(v.inputs[0]) << ", t" << id << ".data, " << slot(id) << ", "
                         << count(id) << ", t" << id << ".count, s);\n"
                         << "            rt.count(\"fir\");\n";
                }
                break;
            case Op::Fft:
                pad(v);
                out_ << "            spectrum_device(plan" << id << ", " << slot(v.inputs[0]) << ", "
                     << slot(id) << ", s);\n"
                     << "            rt.count(\"spectrum\");\n";
                break;
            case Op::Magnitude:
                if (v.spectrum_of >= 0) {
                    out_ << "            // one call: the spectrum %" << v.spectrum_of
                         << " is never stored\n";
                    pad(v);
                    out_ << "            spectrum_magnitude_device(plan" << id << ", "
                         << slot(v.inputs[0]) << ", " << slot(id) << ", s);\n"
                         << "            rt.count(\"spectrum_magnitude\");\n";
                } else {
                    out_ << "            magnitude_device(" << slot(v.inputs[0]) << ", " << slot(id)
                         << ", " << count(id) << ", s);\n"
                         << "            rt.count(\"magnitude\");\n";
                }
                break;
            case Op::Xcorr:
                if (opt_.kernels.xcorr == "fft") {
                    out_ << "            xcorr" << id << ".run(" << slot(v.inputs[0]) << ", "
                         << slot(v.inputs[1]) << ", " << slot(id) << ", s);\n"
                         << "            rt.count(\"xcorr_fft\");\n";
                } else {
                    out_ << "            xcorr_" << opt_.kernels.xcorr << "_device("
                         << slot(v.inputs[0]) << ", " << slot(v.inputs[1]) << ", " << slot(id)
                         << ", " << count(v.inputs[0]) << ", " << count(v.inputs[1]) << ", s);\n"
                         << "            rt.count(\"xcorr\");\n";
                }
                break;
            case Op::Channelize:
                out_ << "            channelize_device(plan" << id << ", " << slot(v.inputs[0])
                     << ", t" << id << ".data, " << slot(id) << ", " << count(v.inputs[0]) << ", "
                     << v.n << ", t" << id << ".count, s);\n"
                     << "            rt.count(\"channelize\");\n";
                break;
            case Op::GccPhat:
                out_ << "            gcc" << id << ".run(" << slot(v.inputs[0]) << ", "
                     << slot(v.inputs[1]) << ", " << slot(id) << ", s);\n"
                     << "            rt.count(\"gcc_phat\");\n";
                break;
            case Op::Capture:
            case Op::Emit:
                break;
            }
        }
        out_ << "        },\n        [&] {\n";
        for (int id : emits_) {
            const Node& value = node(node(id).inputs[0]);
            if (value.op != Op::Capture) {
                out_ << "            rt.download(out" << id << ", " << slot_index(value.id) << ");\n";
            }
        }
        out_ << "        });\n";
    }

    // Zero the FFT input's last-frame padding (the slot may hold an older, longer value).
    void pad(const Node& v)
    {
        const int in = v.inputs[0];
        out_ << "            rt.zero_tail(" << slot_index(in) << ", " << n(in) << " * sizeof(cf32), "
             << n(v.id) << " * sizeof(cf32));\n";
    }

    // A fusion group's FIRs are consecutive in the schedule: the first one makes the batched
    // call for all of them.
    void fused_firs(const Node& v)
    {
        if (!groups_done_.insert(v.group).second) {
            out_ << "            // (in the batched call above)\n";
            return;
        }
        std::vector<int> members;
        for (int id : s_.order) {
            if (node(id).group == v.group) {
                members.push_back(id);
            }
        }
        std::string x, y, counts;
        for (int id : members) {
            const std::string sep = x.empty() ? "" : ", ";
            x += sep + slot(node(id).inputs[0]);
            y += sep + slot(id);
            counts += sep + count(id);
        }
        out_ << "            {  // fusion group " << v.group << ": " << members.size()
             << " FIRs, one launch\n"
             << "                const cf32* x[] = {" << x << "};\n"
             << "                cf32* y[] = {" << y << "};\n"
             << "                const int n[] = {" << counts << "};\n"
             << "                fir_batched_device(" << fir_enum(opt_.kernels.fir) << ", x, y, n, "
             << members.size() << ", t" << v.id << ".data, t" << v.id << ".count, s);\n"
             << "                rt.count(\"fir_batched\");\n"
             << "            }\n";
    }

    void finish()
    {
        out_ << "\n";
        for (int id : emits_) {
            const Node& value = node(node(id).inputs[0]);
            const std::string data = value.op == Op::Capture ? "h" + std::to_string(value.id)
                                                             : "out" + std::to_string(id);
            out_ << "    rt." << (value.type.real ? "write_f32" : "write_cf32") << "("
                 << quote(node(id).path) << ", " << data << ");\n";
        }
        out_ << "    rt.report();\n    return 0;\n";
    }

    const Graph& g_;
    const Schedule& s_;
    const Liveness& l_;
    const HipOptions& opt_;
    std::ostream& out_;
    std::vector<int> captures_, values_, emits_;
    std::set<int> groups_done_;
};

}  // namespace

void emit_hip(Graph g, const HipOptions& opt, const std::string& program_name, std::ostream& out)
{
    if (opt.fuse) {
        fuse_spectrum(g);
    }
    const Schedule s = schedule(g);
    // Downloads follow the whole compute phase, so emitted values keep their slots to the end.
    const Liveness l = liveness(g, s, opt.reuse, true);
    HipEmitter(g, s, l, opt, out).run(program_name);
}

}  // namespace sigflow
