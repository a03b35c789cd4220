// sigflowc: the SigFlow pipeline language (compiler/README.md of the reference) compiled to HIP
// programs over this library's device-pointer operator surface (host/device_ops.hpp).
//
// Stages and the data each hands to the next:
//   lex / parse    Source -> tokens -> Program (syntax tree)               syntax.cpp
//   analyze        names, arguments, types, sample rates -> Graph          sema.cpp
//   plan           dead nodes, FIR fusion groups, launch order, slots      passes.cpp
//   print_* / emit_hip                                                     output.cpp, hip.cpp
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

namespace sigflow {

// --- source and diagnostics -------------------------------------------------------------------

struct Pos {
    int line = 1;
    int col = 1;  // 1-based
    int len = 1;  // characters a diagnostic underlines
};

class Source {
public:
    Source(std::string name, std::string text);
    const std::string& name() const { return name_; }
    const std::string& text() const { return text_; }
    std::string line(int number) const;  // without its line break

private:
    std::string name_;
    std::string text_;
    std::vector<std::size_t> starts_;
};

// "file:line:col: error: message", then the source line and a caret under the position.
// Diagnostics print sorted by position, errors and warnings interleaved.
class Diagnostics {
public:
    explicit Diagnostics(const Source& src) : src_(src) {}
    void error(const Pos& at, const std::string& message) { add(true, at, message); }
    void warning(const Pos& at, const std::string& message) { add(false, at, message); }
    bool failed() const { return errors_ > 0; }
    void print(std::ostream& out);  // and forget them

private:
    struct Entry {
        bool error;
        Pos at;
        std::string message;
    };
    void add(bool error, const Pos& at, const std::string& message);
    const Source& src_;
    std::vector<Entry> entries_;
    int errors_ = 0;
};

// Thrown to stop a stage once its diagnostics are recorded.
struct Failed {};

// --- syntax -----------------------------------------------------------------------------------

struct Token {
    enum class Kind { Ident, Emit, Number, String, LParen, RParen, Comma, Equals, Pipe, Arrow,
                      Newline, End };
    Kind kind;
    std::string text;  // string tokens: the contents, without quotes
    Pos pos;
};
const char* token_kind_name(Token::Kind kind);

struct Expr;
using ExprPtr = std::unique_ptr<Expr>;

struct Kwarg {
    std::string name;
    Pos name_pos;
    ExprPtr value;  // a Number or String literal
};

struct Expr {
    enum class Kind { Name, Number, String, Call, Pipe } kind = Kind::Name;
    std::string text;             // name, callee, string contents, number spelling
    double number = 0.0;
    Pos pos;                      // Call: the callee; Pipe: its input's; literals with quotes
    std::vector<ExprPtr> args;    // Call: positional arguments
    std::vector<Kwarg> kwargs;    // Call
    bool bare = false;            // Call written without parentheses (only after '|>')
    ExprPtr input;                // Pipe: the value piped in
    ExprPtr call;                 // Pipe: the Call it feeds
};

struct Stmt {
    bool is_emit = false;
    std::string name;  // bound name, or the emitted name
    Pos name_pos;
    std::string path;  // emit target
    Pos path_pos;
    ExprPtr value;     // binding
    Pos pos;           // the statement's first token
};

struct Program {
    std::vector<Stmt> stmts;
};

std::vector<Token> lex(const Source& src, Diagnostics& diag);
Program parse(const std::vector<Token>& tokens, Diagnostics& diag);

// --- dataflow graph ---------------------------------------------------------------------------

enum class Op { Capture, Fir, Fft, Magnitude, Xcorr, Channelize, GccPhat, Emit };
const char* op_name(Op op);

struct Type {
    bool frames = false;      // frames<elem, n>; else stream<cf32>
    bool real = false;        // element f32 (magnitude of frames); else cf32
    long long n = 0;          // frame length
    double rate = 0.0;        // samples per second
    std::string rate_text;    // as the capture spelled it
};
std::string type_name(const Type& t);  // "stream<cf32>", "frames<f32, 1024>"

struct Node {
    int id = 0;               // index in Graph::nodes
    Op op = Op::Capture;
    std::vector<int> inputs;
    Type type;                // unused for Emit
    std::string name;         // binding the node defines ("" for intermediate values)
    std::string path;         // Capture: read; Emit: written
    std::string taps;         // Fir, Channelize
    long long n = 0;          // Fft: frame length; Channelize: channels
    int group = -1;           // FIR fusion group
    int spectrum_of = -1;     // Magnitude only: the fft node fused into it (HIP backend)
    Pos pos;                  // callee (or emit statement) position: the source-order key
};

struct Graph {
    std::vector<Node> nodes;  // inputs have smaller ids than their users
    std::vector<bool> live;
    bool before(int a, int b) const;  // source order, then id
};

// Semantic analysis and graph construction. Throws Failed after recording errors. Ids follow
// first use: statements are visited in order and every name is built, inputs first, the first
// time it is reached.
Graph analyze(const Program& prog, Diagnostics& diag);

// --- passes -----------------------------------------------------------------------------------

constexpr int kMaxFusedFirs = 8;  // FIRs in one batched launch (the library's RF_MAX_FIR_BATCH)

void eliminate_dead_nodes(Graph& g);
void fuse_firs(Graph& g);

struct Schedule {
    std::vector<int> order;   // node ids, emits included
    std::vector<int> launch;  // launch index of each step (a fusion group is one launch)
    std::vector<int> stuck;   // live nodes that never became ready (a cycle)
};
Schedule schedule(const Graph& g);

struct LiveRange {
    int node = 0;
    int def = 0;    // step that writes the value
    int last = 0;   // last step that reads it (== def when nothing does)
    int slot = 0;
};
struct Liveness {
    std::vector<LiveRange> ranges;  // value nodes in schedule order
    std::vector<int> slot_of;       // by node id; -1 for emits and dead nodes
    int slots = 0;
};
// With `hold_emitted`, an emitted value stays live to the last step: for a backend that copies
// every emitted value back after the whole compute phase (the HIP backend does), not at the
// emit's own step. Without it (--emit=liveness, json), an emit is a reader like any other.
Liveness liveness(const Graph& g, const Schedule& s, bool reuse, bool hold_emitted = false);

// Bytes of every value (-1 unknown): capture sizes from files under `input_dir`.
std::vector<long long> value_bytes(const Graph& g, const std::string& input_dir);

struct MemoryPlan {
    long long no_reuse = -1;
    long long reuse = -1;
    long long live_peak = -1;
};
MemoryPlan plan_memory(const Liveness& l, const std::vector<long long>& bytes);

// --- output -----------------------------------------------------------------------------------

struct Kernels {
    std::string fir = "naive";    // naive | tiled | tiled_const | register | fft_ola
    std::string xcorr = "naive";  // naive | tiled | fft
};
const std::vector<std::string>& fir_impl_names();
const std::vector<std::string>& xcorr_impl_names();

void print_tokens(const std::vector<Token>& tokens, std::ostream& out);
void print_ast(const Program& prog, std::ostream& out);
std::string node_text(const Node& node);  // one --emit=ir line, no line break
void print_ir(const Graph& g, std::ostream& out);
void print_schedule(const Graph& g, const Schedule& s, std::ostream& out);
void print_liveness(const Graph& g, const Liveness& l, const std::vector<long long>& bytes,
                    const std::string& input_dir, std::ostream& out);
void print_json(const Graph& g, bool dce, bool fuse, const Schedule& s, const Liveness& l,
                const std::vector<long long>& bytes, std::ostream& out);

// The HIP backend: a C++ program over host/device_ops.hpp and host/pipeline_runtime.hpp that
// keeps every value in GPU memory and runs all work on one stream (optionally as one hipGraph).
// With `fuse`, an fft whose only reader is a magnitude becomes one spectrum-magnitude call
// (fuse_spectrum); the backend then schedules and assigns slots again on that graph.
struct HipOptions {
    Kernels kernels;
    bool reuse = true;
    bool fuse = true;
};
void fuse_spectrum(Graph& g);
void emit_hip(Graph g, const HipOptions& opt, const std::string& program_name, std::ostream& out);

}  // namespace sigflow
