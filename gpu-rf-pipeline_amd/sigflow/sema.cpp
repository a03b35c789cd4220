// Semantic analysis and graph construction for SigFlow.
//
// Checks, with the diagnostics the reference's compiler prints for them (its golden files
// compiler/tests/golden/sema_*.expected): names (unknown, redefined, an operation used as a
// value or a value called), cycles through bindings, positional and keyword arguments, value
// types (stream<cf32>, frames<cf32, N>, frames<f32, N>), literal values (rate, n, channels),
// equal sample rates for two-input operations, duplicate emit paths, and the file extension of
// emitted values (a warning).
//
// The graph is built only from a program with no errors. Nodes are created depth first, inputs
// before users, visiting the statements in order, so ids follow first use; an alias (`d = c`)
// makes no node.

#include "sigflow.hpp"

#include <cmath>
#include <optional>

namespace sigflow {

std::string type_name(const Type& t)
{
    if (!t.frames) {
        return "stream<cf32>";
    }
    return std::string("frames<") + (t.real ? "f32" : "cf32") + ", " + std::to_string(t.n) + ">";
}

const char* op_name(Op op)
{
    switch (op) {
    case Op::Capture: return "capture";
    case Op::Fir: return "fir";
    case Op::Fft: return "fft";
    case Op::Magnitude: return "magnitude";
    case Op::Xcorr: return "xcorr";
    case Op::Channelize: return "channelize";
    case Op::GccPhat: return "gcc_phat";
    case Op::Emit: return "emit";
    }
    return "?";
}

bool Graph::before(int a, int b) const
{
    const Pos& pa = nodes[static_cast<std::size_t>(a)].pos;
    const Pos& pb = nodes[static_cast<std::size_t>(b)].pos;
    if (pa.line != pb.line) {
        return pa.line < pb.line;
    }
    if (pa.col != pb.col) {
        return pa.col < pb.col;
    }
    return a < b;
}

namespace {

// Keyword arguments: name and whether it takes a number (else a string). All are required.
struct KwSpec {
    const char* name;
    bool number;
};

struct OpSpec {
    const char* name;
    Op op;
    int positional;
    std::vector<KwSpec> kwargs;  // sorted by name
};

// Sorted by name: the order of the "available" list.
const std::vector<OpSpec>& ops()
{
    static const std::vector<OpSpec> table = {
        {"capture", Op::Capture, 1, {{"rate", true}}},
        {"channelize", Op::Channelize, 1, {{"channels", true}, {"taps", false}}},
        {"fft", Op::Fft, 1, {{"n", true}}},
        {"fir", Op::Fir, 1, {{"taps", false}}},
        {"gcc_phat", Op::GccPhat, 2, {}},
        {"magnitude", Op::Magnitude, 1, {}},
        {"xcorr", Op::Xcorr, 2, {}},
    };
    return table;
}

const OpSpec* find_op(const std::string& name)
{
    for (const OpSpec& s : ops()) {
        if (name == s.name) {
            return &s;
        }
    }
    return nullptr;
}

std::string op_list()
{
    std::string out;
    for (const OpSpec& s : ops()) {
        out += (out.empty() ? "" : ", ") + std::string(s.name);
    }
    return out;
}

constexpr long long kMaxFft = 1LL << 24;      // the library's longest transform
// Depth of bindings resolved through one another (each level is a few stack frames here and in
// the graph builder).
constexpr std::size_t kMaxBindingDepth = 1000;
constexpr long long kMaxChannels = 1LL << 20;

class Analyzer {
public:
    Analyzer(const Program& p, Diagnostics& d) : prog_(p), diag_(d) {}

    Graph run()
    {
        collect_names();
        bool any_emit = false;
        std::map<std::string, int> written;  // emit path -> line
        for (std::size_t i = 0; i < prog_.stmts.size(); ++i) {
            const Stmt& st = prog_.stmts[i];
            if (!st.is_emit) {
                if (first_def_.at(st.name) == i) {
                    binding_type(st.name, st.name_pos);
                } else {
                    expr_type(*st.value, true);  // a redefinition is still checked
                }
                continue;
            }
            any_emit = true;
            const std::optional<Type> t = name_type(st.name, st.name_pos);
            const auto seen = written.find(st.path);
            if (seen != written.end()) {
                diag_.error(st.path_pos, "\"" + st.path + "\" is already written by the emit on line "
                                             + std::to_string(seen->second));
            } else {
                written.emplace(st.path, st.pos.line);
            }
            if (t) {
                const std::string ext = t->real ? ".f32" : ".cf32";
                const bool ok = st.path.size() >= ext.size()
                                && st.path.compare(st.path.size() - ext.size(), ext.size(), ext) == 0;
                if (!ok) {
                    diag_.warning(st.path_pos, "'" + st.name + "' is " + type_name(*t)
                                                   + ", which is written as raw " + ext.substr(1)
                                                   + "; the path does not end in " + ext);
                }
            }
        }
        if (!any_emit) {
            diag_.error(Pos{1, 1, 1},
                        "program has no 'emit' statements, so there is nothing to compute");
        }
        if (diag_.failed()) {
            throw Failed{};
        }
        return build();
    }

private:
    // --- names ---------------------------------------------------------------------------------

    void collect_names()
    {
        for (std::size_t i = 0; i < prog_.stmts.size(); ++i) {
            const Stmt& st = prog_.stmts[i];
            if (st.is_emit) {
                continue;
            }
            const auto it = first_def_.find(st.name);
            if (it != first_def_.end()) {
                diag_.error(st.name_pos, "redefinition of '" + st.name + "' (first defined on line "
                                             + std::to_string(prog_.stmts[it->second].pos.line)
                                             + ")");
                continue;
            }
            first_def_.emplace(st.name, i);
        }
    }

    // Type of a name used as a value; nullopt after an error (reported once, where it is).
    std::optional<Type> name_type(const std::string& name, const Pos& at)
    {
        if (first_def_.count(name) == 0) {
            if (find_op(name) != nullptr) {
                diag_.error(at, "'" + name + "' is an operation, not a binding; call it as " + name
                                    + "(...) or after '|>'");
            } else {
                diag_.error(at, "unknown name '" + name + "'");
            }
            return std::nullopt;
        }
        return binding_type(name, at);
    }

    std::optional<Type> binding_type(const std::string& name, const Pos& at)
    {
        const auto done = types_.find(name);
        if (done != types_.end()) {
            return done->second;
        }
        const auto open = on_stack_.find(name);
        if (open != on_stack_.end()) {
            std::string path;
            for (std::size_t j = open->second; j < stack_.size(); ++j) {
                path += stack_[j] + " -> ";
            }
            diag_.error(at, "cycle: '" + name + "' depends on itself (" + path + name + ")");
            return std::nullopt;
        }
        if (stack_.size() >= kMaxBindingDepth) {
            diag_.error(at, "bindings nested too deeply (more than "
                                + std::to_string(kMaxBindingDepth) + " levels through '" + name
                                + "')");
            types_.emplace(name, std::nullopt);
            return std::nullopt;
        }
        on_stack_.emplace(name, stack_.size());
        stack_.push_back(name);
        const Stmt& st = prog_.stmts[first_def_.at(name)];
        std::optional<Type> t = expr_type(*st.value, true);
        stack_.pop_back();
        on_stack_.erase(name);
        types_.emplace(name, t);
        return t;
    }

    // --- expressions ---------------------------------------------------------------------------

    std::optional<Type> expr_type(const Expr& e, bool bound)
    {
        switch (e.kind) {
        case Expr::Kind::Name:
            return name_type(e.text, e.pos);
        case Expr::Kind::Number:
        case Expr::Kind::String:
            if (bound) {
                diag_.error(e.pos, std::string("expected a signal, found a ")
                                       + (e.kind == Expr::Kind::Number ? "number" : "string")
                                       + " literal");
            }
            return std::nullopt;
        case Expr::Kind::Call:
            return remember(e, call_type(e, nullptr));
        case Expr::Kind::Pipe:
            return remember(e, call_type(*e.call, e.input.get()));
        }
        return std::nullopt;
    }

    std::optional<Type> remember(const Expr& e, std::optional<Type> t)
    {
        if (t) {
            expr_types_[&e] = *t;
        }
        return t;
    }

    std::optional<Type> call_type(const Expr& call, const Expr* piped)
    {
        const OpSpec* spec = find_op(call.text);
        if (spec == nullptr) {
            if (first_def_.count(call.text) != 0) {
                diag_.error(call.pos, "'" + call.text + "' is a binding, not an operation");
            } else {
                diag_.error(call.pos, "unknown operation '" + call.text + "' (available: "
                                          + op_list() + ")");
            }
            return std::nullopt;
        }
        std::vector<const Expr*> args;
        if (piped != nullptr) {
            args.push_back(piped);
        }
        for (const ExprPtr& a : call.args) {
            args.push_back(a.get());
        }
        if (static_cast<int>(args.size()) != spec->positional) {
            diag_.error(call.pos, call.text + " expects " + std::to_string(spec->positional)
                                      + " positional argument" + (spec->positional == 1 ? "" : "s")
                                      + ", got " + std::to_string(args.size())
                                      + (piped != nullptr
                                             ? " (counting the input piped in with '|>')"
                                             : ""));
            return std::nullopt;
        }

        // Keyword arguments: known, not repeated, of the right literal kind, all present.
        bool ok = true;
        std::map<std::string, const Kwarg*> given;
        for (const Kwarg& k : call.kwargs) {
            const KwSpec* ks = nullptr;
            for (const KwSpec& s : spec->kwargs) {
                if (k.name == s.name) {
                    ks = &s;
                }
            }
            if (ks == nullptr) {
                std::string known;
                for (const KwSpec& s : spec->kwargs) {
                    known += (known.empty() ? "" : ", ") + std::string(s.name);
                }
                diag_.error(k.name_pos, call.text + " has no keyword argument '" + k.name + "' ("
                                            + (known.empty() ? "it takes none"
                                                             : "expected: " + known)
                                            + ")");
                ok = false;
                continue;
            }
            if (given.count(k.name) != 0) {
                diag_.error(k.name_pos, "duplicate keyword argument '" + k.name + "'");
                ok = false;
                continue;
            }
            const bool is_number = k.value->kind == Expr::Kind::Number;
            if (is_number != ks->number) {
                diag_.error(k.value->pos, "keyword argument '" + k.name + "' of " + call.text
                                              + " must be a " + (ks->number ? "number" : "string")
                                              + ", got a " + (is_number ? "number" : "string"));
                ok = false;
                continue;
            }
            given.emplace(k.name, &k);
        }
        for (const KwSpec& s : spec->kwargs) {
            if (given.count(s.name) == 0 && ok) {
                diag_.error(call.pos, call.text + " requires keyword argument '" + s.name + "'");
                ok = false;
            }
        }
        if (!ok) {
            return std::nullopt;
        }

        if (spec->op == Op::Capture) {
            if (args[0]->kind != Expr::Kind::String) {
                diag_.error(args[0]->pos, "capture expects a string literal path as argument 1");
                ok = false;
            }
            const Expr& rate = *given.at("rate")->value;
            if (!(rate.number > 0.0) || !std::isfinite(rate.number)) {
                diag_.error(rate.pos, "sample rate must be a positive number, got " + rate.text);
                ok = false;
            }
            if (!ok) {
                return std::nullopt;
            }
            Type t;
            t.rate = rate.number;
            t.rate_text = rate.text;
            return t;
        }

        // Signal inputs: evaluated in order, every one of them even after an error.
        std::vector<std::optional<Type>> in;
        for (std::size_t i = 0; i < args.size(); ++i) {
            const Expr& a = *args[i];
            if (a.kind == Expr::Kind::Number || a.kind == Expr::Kind::String) {
                diag_.error(a.pos, call.text + " expects a signal as argument " + std::to_string(i + 1)
                                       + ", got a "
                                       + (a.kind == Expr::Kind::Number ? "number" : "string")
                                       + " literal");
                in.push_back(std::nullopt);
                continue;
            }
            in.push_back(expr_type(a, false));
        }

        Type out;
        switch (spec->op) {
        case Op::Fir:
        case Op::Fft:
        case Op::Channelize:
            if (!stream_input(call, args, in, 0)) {
                return std::nullopt;
            }
            out = *in[0];
            if (spec->op == Op::Fft) {
                const Expr& n = *given.at("n")->value;
                if (!(n.number >= 1.0) || n.number != std::floor(n.number)) {
                    diag_.error(n.pos, "fft size n must be a positive integer, got " + n.text);
                    return std::nullopt;
                }
                if (n.number > static_cast<double>(kMaxFft)) {
                    diag_.error(n.pos, "fft size n must be at most " + std::to_string(kMaxFft)
                                           + ", got " + n.text);
                    return std::nullopt;
                }
                out.frames = true;
                out.n = static_cast<long long>(n.number);
            } else if (spec->op == Op::Channelize) {
                const Expr& c = *given.at("channels")->value;
                if (!(c.number >= 1.0) || c.number != std::floor(c.number)
                    || c.number > static_cast<double>(kMaxChannels)) {
                    diag_.error(c.pos, "channels must be an integer from 1 to "
                                           + std::to_string(kMaxChannels) + ", got " + c.text);
                    return std::nullopt;
                }
                out.frames = true;
                out.n = static_cast<long long>(c.number);
            }
            return out;
        case Op::Magnitude:
            if (!in[0]) {
                return std::nullopt;
            }
            if (!in[0]->frames || in[0]->real) {
                diag_.error(args[0]->pos,
                            "magnitude expects frames<cf32, N>, got " + type_name(*in[0])
                                + (in[0]->frames ? "" : "; apply fft(n=...) first"));
                return std::nullopt;
            }
            out = *in[0];
            out.real = true;
            return out;
        case Op::Xcorr:
        case Op::GccPhat: {
            const bool a = stream_input(call, args, in, 0);
            const bool b = stream_input(call, args, in, 1);
            if (!a || !b) {
                return std::nullopt;
            }
            if (in[0]->rate != in[1]->rate) {
                diag_.error(call.pos, call.text + " inputs have different sample rates ("
                                          + in[0]->rate_text + " vs " + in[1]->rate_text + ")");
                return std::nullopt;
            }
            return *in[0];
        }
        default:
            return std::nullopt;
        }
    }

    bool stream_input(const Expr& call, const std::vector<const Expr*>& args,
                      const std::vector<std::optional<Type>>& in, std::size_t i)
    {
        if (!in[i]) {
            return false;
        }
        if (in[i]->frames) {
            diag_.error(args[i]->pos, call.text + " expects stream<cf32> as argument "
                                          + std::to_string(i + 1) + ", got " + type_name(*in[i]));
            return false;
        }
        return true;
    }

    // --- graph ---------------------------------------------------------------------------------

    Graph build()
    {
        for (const Stmt& st : prog_.stmts) {
            if (!st.is_emit) {
                binding_node(st.name);
                continue;
            }
            Node emit;
            emit.op = Op::Emit;
            emit.inputs = {binding_node(st.name)};
            emit.path = st.path;
            emit.pos = st.pos;
            add(std::move(emit));
        }
        g_.live.assign(g_.nodes.size(), true);
        return std::move(g_);
    }

    int add(Node node)
    {
        node.id = static_cast<int>(g_.nodes.size());
        g_.nodes.push_back(std::move(node));
        return g_.nodes.back().id;
    }

    int binding_node(const std::string& name)
    {
        const auto done = nodes_.find(name);
        if (done != nodes_.end()) {
            return done->second;
        }
        const int id = expr_node(*prog_.stmts[first_def_.at(name)].value);
        Node& n = g_.nodes[static_cast<std::size_t>(id)];
        if (n.name.empty()) {
            n.name = name;
        }
        nodes_.emplace(name, id);
        return id;
    }

    int expr_node(const Expr& e)
    {
        if (e.kind == Expr::Kind::Name) {
            return binding_node(e.text);
        }
        const bool pipe = e.kind == Expr::Kind::Pipe;
        const Expr& call = pipe ? *e.call : e;
        Node node;
        node.op = find_op(call.text)->op;
        node.type = expr_types_.at(&e);
        node.pos = call.pos;
        std::vector<const Expr*> args;
        if (pipe) {
            args.push_back(e.input.get());
        }
        for (const ExprPtr& a : call.args) {
            args.push_back(a.get());
        }
        if (node.op == Op::Capture) {
            node.path = args[0]->text;
        } else {
            for (const Expr* a : args) {
                node.inputs.push_back(expr_node(*a));
            }
        }
        for (const Kwarg& k : call.kwargs) {
            if (k.name == "taps") {
                node.taps = k.value->text;
            } else if (k.name == "n" || k.name == "channels") {
                node.n = static_cast<long long>(k.value->number);
            }
        }
        return add(std::move(node));
    }

    const Program& prog_;
    Diagnostics& diag_;
    std::map<std::string, std::size_t> first_def_;
    std::map<std::string, std::optional<Type>> types_;
    std::map<const Expr*, Type> expr_types_;
    std::vector<std::string> stack_;           // bindings being resolved, outermost first
    std::map<std::string, std::size_t> on_stack_;  // their positions in stack_
    Graph g_;
    std::map<std::string, int> nodes_;
};

}  // namespace

Graph analyze(const Program& prog, Diagnostics& diag)
{
    return Analyzer(prog, diag).run();
}

}  // namespace sigflow
