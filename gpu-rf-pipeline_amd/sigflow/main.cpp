// sigflowc: compile a SigFlow program (.sf) to a HIP program, or print one of its stages.
//
//   sigflowc PROGRAM.sf [--emit=MODE] [-o FILE] [--no-dce] [--no-fuse] [--no-reuse]
//            [--fir-impl=NAME] [--xcorr-impl=NAME] [--input-dir=DIR]
//
// MODE: tokens, ast, check, ir, schedule, liveness, json, hip (the default). Flags and exit
// codes follow the reference's compiler/src/driver.cpp:119-286 (0 success, 1 a compile error,
// 2 a usage error); its script and cuda modes are replaced by hip.

#include "sigflow.hpp"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {

constexpr int kOk = 0;
constexpr int kCompileError = 1;
constexpr int kUsage = 2;

const char* kModes[] = {"tokens", "ast", "check", "ir", "schedule", "liveness", "json", "hip"};

struct Options {
    std::string input;
    std::string output;
    std::string mode = "hip";
    bool dce = true;
    bool fuse = true;
    bool reuse = true;
    sigflow::Kernels kernels;
    std::string input_dir = ".";
};

void usage(std::ostream& out)
{
    out << "usage: sigflowc PROGRAM.sf [--emit=MODE] [-o FILE] [--no-dce] [--no-fuse] "
           "[--no-reuse]\n"
           "                [--fir-impl=NAME] [--xcorr-impl=NAME] [--input-dir=DIR]\n"
           "  MODE: tokens, ast, check, ir, schedule, liveness, json, hip (default)\n";
}

std::string join(const std::vector<std::string>& names)
{
    std::string out;
    for (const std::string& n : names) {
        out += (out.empty() ? "" : ", ") + n;
    }
    return out;
}

// False after printing the first problem.
bool parse_args(int argc, char** argv, Options& o, std::ostream& err)
{
    bool have_input = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-o") {
            if (i + 1 == argc) {
                err << "sigflowc: error: -o needs a file name\n";
                return false;
            }
            o.output = argv[++i];
        } else if (a.rfind("--emit=", 0) == 0) {
            o.mode = a.substr(7);
            if (std::find(std::begin(kModes), std::end(kModes), o.mode) == std::end(kModes)) {
                err << "sigflowc: error: unknown --emit mode '" << o.mode << "' (available: "
                    << join(std::vector<std::string>(std::begin(kModes), std::end(kModes)))
                    << ")\n";
                return false;
            }
        } else if (a == "--no-dce") {
            o.dce = false;
        } else if (a == "--no-fuse") {
            o.fuse = false;
        } else if (a == "--no-reuse") {
            o.reuse = false;
        } else if (a.rfind("--fir-impl=", 0) == 0 || a.rfind("--xcorr-impl=", 0) == 0) {
            const bool fir = a[2] == 'f';
            const std::string value = a.substr(fir ? 11 : 13);
            const std::vector<std::string>& names =
                fir ? sigflow::fir_impl_names() : sigflow::xcorr_impl_names();
            if (std::find(names.begin(), names.end(), value) == names.end()) {
                err << "sigflowc: error: unknown " << (fir ? "--fir-impl" : "--xcorr-impl") << " '"
                    << value << "' (available: " << join(names) << ")\n";
                return false;
            }
            (fir ? o.kernels.fir : o.kernels.xcorr) = value;
        } else if (a.rfind("--input-dir=", 0) == 0) {
            o.input_dir = a.substr(12);
        } else if (a == "--help" || a == "-h") {
            usage(std::cout);
            std::exit(kOk);
        } else if (!a.empty() && a[0] == '-') {
            err << "sigflowc: error: unknown option '" << a << "'\n";
            return false;
        } else if (have_input) {
            err << "sigflowc: error: more than one input file ('" << o.input << "' and '" << a
                << "')\n";
            return false;
        } else {
            o.input = a;
            have_input = true;
        }
    }
    if (!have_input) {
        err << "sigflowc: error: no input file\n";
        return false;
    }
    return true;
}

int compile(const Options& o, std::ostream& out, std::ostream& err)
{
    std::ifstream file(o.input, std::ios::binary);
    if (!file) {
        err << "sigflowc: error: cannot read '" << o.input << "'\n";
        return kUsage;
    }
    std::stringstream text;
    text << file.rdbuf();
    const sigflow::Source src(o.input, text.str());
    sigflow::Diagnostics diag(src);
    const auto stop = [&] {
        diag.print(err);
        err.flush();
        return kCompileError;
    };

    const std::vector<sigflow::Token> tokens = sigflow::lex(src, diag);
    if (diag.failed()) {
        return stop();
    }
    if (o.mode == "tokens") {
        diag.print(err);
        err.flush();
        sigflow::print_tokens(tokens, out);
        return kOk;
    }
    const sigflow::Program prog = sigflow::parse(tokens, diag);
    if (diag.failed()) {
        return stop();
    }
    if (o.mode == "ast") {
        diag.print(err);
        err.flush();
        sigflow::print_ast(prog, out);
        return kOk;
    }
    sigflow::Graph g;
    try {
        g = sigflow::analyze(prog, diag);
    } catch (const sigflow::Failed&) {
        return stop();
    }
    diag.print(err);
    err.flush();
    if (o.mode == "check") {
        return kOk;
    }

    if (o.dce) {
        sigflow::eliminate_dead_nodes(g);
    }
    if (o.fuse) {
        sigflow::fuse_firs(g);
    }
    const sigflow::Schedule s = sigflow::schedule(g);
    if (!s.stuck.empty()) {
        // analyze() rejects cycles, so this is a bug of the compiler, reported as a diagnostic.
        std::string ids;
        for (int id : s.stuck) {
            ids += (ids.empty() ? "%" : ", %") + std::to_string(id);
        }
        diag.error(g.nodes[static_cast<std::size_t>(s.stuck[0])].pos,
                   "internal error: the IR has a cycle through " + ids);
        return stop();
    }
    const sigflow::Liveness l = sigflow::liveness(g, s, o.reuse);

    if (o.mode == "ir") {
        sigflow::print_ir(g, out);
    } else if (o.mode == "schedule") {
        sigflow::print_schedule(g, s, out);
    } else if (o.mode == "liveness") {
        sigflow::print_liveness(g, l, sigflow::value_bytes(g, o.input_dir), o.input_dir, out);
    } else if (o.mode == "json") {
        sigflow::print_json(g, o.dce, o.fuse, s, l, sigflow::value_bytes(g, o.input_dir), out);
    } else {
        sigflow::HipOptions hip;
        hip.kernels = o.kernels;
        hip.reuse = o.reuse;
        hip.fuse = o.fuse;
        const std::size_t slash = o.input.find_last_of('/');
        sigflow::emit_hip(g, hip, slash == std::string::npos ? o.input : o.input.substr(slash + 1),
                          out);
    }
    return kOk;
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    if (!parse_args(argc, argv, o, std::cerr)) {
        std::cerr << "sigflowc: run 'sigflowc --help' for usage\n";
        return kUsage;
    }
    if (o.output.empty()) {
        return compile(o, std::cout, std::cerr);
    }
    std::ostringstream text;
    const int code = compile(o, text, std::cerr);
    if (code == kOk) {
        std::ofstream file(o.output, std::ios::binary);
        file << text.str();
        if (!file) {
            std::cerr << "sigflowc: error: cannot write '" << o.output << "'\n";
            return kUsage;
        }
    }
    return code;
}
