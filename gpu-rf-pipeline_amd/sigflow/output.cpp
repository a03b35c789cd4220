// Text forms of sigflowc's stages: --emit=tokens, ast, ir, schedule, liveness and json. Their
// layouts are the reference compiler's (compiler/src/driver.cpp print_tokens, ast.cpp
// print_ast, ir_builder.cpp format_node, liveness.cpp print_liveness, json.cpp print_json),
// which its golden files pin byte for byte and the reference's viewer and e2e tests read.

#include "sigflow.hpp"

namespace sigflow {

void print_tokens(const std::vector<Token>& tokens, std::ostream& out)
{
    for (const Token& t : tokens) {
        out << t.pos.line << ":" << t.pos.col << " " << token_kind_name(t.kind);
        if (t.kind == Token::Kind::String) {
            out << " \"" << t.text << "\"";
        } else if (t.kind != Token::Kind::Newline && t.kind != Token::Kind::End) {
            out << " " << t.text;
        }
        out << "\n";
    }
}

namespace {

std::string at(const Pos& p)
{
    return " @" + std::to_string(p.line) + ":" + std::to_string(p.col);
}

std::string literal(const Expr& e)
{
    return e.kind == Expr::Kind::Number ? "number " + e.text : "string \"" + e.text + "\"";
}

void print_expr(const Expr& e, int depth, std::ostream& out)
{
    const std::string indent(static_cast<std::size_t>(depth) * 2, ' ');
    switch (e.kind) {
    case Expr::Kind::Name:
        out << indent << "name " << e.text << at(e.pos) << "\n";
        return;
    case Expr::Kind::Number:
    case Expr::Kind::String:
        out << indent << literal(e) << at(e.pos) << "\n";
        return;
    case Expr::Kind::Call:
        out << indent << "call " << e.text << (e.bare ? " (bare)" : "") << at(e.pos) << "\n";
        for (const ExprPtr& a : e.args) {
            print_expr(*a, depth + 1, out);
        }
        for (const Kwarg& k : e.kwargs) {
            out << indent << "  kwarg " << k.name << " = " << literal(*k.value) << at(k.name_pos)
                << "\n";
        }
        return;
    case Expr::Kind::Pipe:
        out << indent << "pipe" << at(e.pos) << "\n";
        print_expr(*e.input, depth + 1, out);
        print_expr(*e.call, depth + 1, out);
        return;
    }
}

std::string json_string(const std::string& text)
{
    std::string out = "\"";
    for (char c : text) {
        if (c == '"' || c == '\\') {
            out += '\\';
            out += c;
        } else if (static_cast<unsigned char>(c) < 0x20) {
            static const char* hex = "0123456789abcdef";
            out += "\\u00";
            out += hex[(c >> 4) & 0xf];
            out += hex[c & 0xf];
        } else {
            out += c;
        }
    }
    return out + "\"";
}

}  // namespace

void print_ast(const Program& prog, std::ostream& out)
{
    out << "program\n";
    for (const Stmt& st : prog.stmts) {
        if (st.is_emit) {
            out << "  emit " << st.name << " -> \"" << st.path << "\"" << at(st.pos) << "\n";
            continue;
        }
        out << "  binding " << st.name << at(st.pos) << "\n";
        print_expr(*st.value, 2, out);
    }
}

std::string node_text(const Node& n)
{
    if (n.op == Op::Emit) {
        return "emit %" + std::to_string(n.inputs[0]) + " -> \"" + n.path + "\"";
    }
    std::string s = "%" + std::to_string(n.id) + " = " + op_name(n.op);
    for (int in : n.inputs) {
        s += " %" + std::to_string(in);
    }
    if (n.op == Op::Capture) {
        s += " \"" + n.path + "\"";
    }
    if (n.op == Op::Fir || n.op == Op::Channelize) {
        s += " taps=\"" + n.taps + "\"";
    }
    if (n.op == Op::Channelize) {
        s += " channels=" + std::to_string(n.n);
    }
    if (n.op == Op::Fft) {
        s += " n=" + std::to_string(n.n);
    }
    s += " : " + type_name(n.type) + " @" + n.type.rate_text;
    if (n.group >= 0) {
        s += " [fused " + std::to_string(n.group) + "]";
    }
    if (!n.name.empty()) {
        s += "   # " + n.name;
    }
    return s;
}

void print_ir(const Graph& g, std::ostream& out)
{
    for (const Node& n : g.nodes) {
        if (g.live[static_cast<std::size_t>(n.id)]) {
            out << node_text(n) << "\n";
        }
    }
}

void print_schedule(const Graph& g, const Schedule& s, std::ostream& out)
{
    for (std::size_t k = 0; k < s.order.size(); ++k) {
        out << k << ": " << node_text(g.nodes[static_cast<std::size_t>(s.order[k])]) << "\n";
    }
}

void print_liveness(const Graph& g, const Liveness& l, const std::vector<long long>& bytes,
                    const std::string& input_dir, std::ostream& out)
{
    for (const LiveRange& r : l.ranges) {
        const Node& n = g.nodes[static_cast<std::size_t>(r.node)];
        const long long b = bytes[static_cast<std::size_t>(r.node)];
        out << "%" << r.node << " " << op_name(n.op) << ": live " << r.def << ".." << r.last
            << ", slot " << r.slot << ", " << (b < 0 ? std::string("?") : std::to_string(b))
            << " bytes";
        if (!n.name.empty()) {
            out << "   # " << n.name;
        }
        out << "\n";
    }
    out << l.slots << " slots for " << l.ranges.size() << " values\n";
    const MemoryPlan plan = plan_memory(l, bytes);
    if (plan.reuse < 0) {
        out << "planned GPU memory: unknown (a capture file is missing under '" << input_dir
            << "'; see --input-dir)\n";
        return;
    }
    out << "planned GPU memory: " << plan.no_reuse << " bytes without reuse, " << plan.reuse
        << " with reuse, " << plan.live_peak << " live at the busiest step\n";
}

void print_json(const Graph& g, bool dce, bool fuse, const Schedule& s, const Liveness& l,
                const std::vector<long long>& bytes, std::ostream& out)
{
    std::vector<int> step(g.nodes.size(), -1);
    for (std::size_t k = 0; k < s.order.size(); ++k) {
        step[static_cast<std::size_t>(s.order[k])] = static_cast<int>(k);
    }
    std::vector<const LiveRange*> range(g.nodes.size(), nullptr);
    for (const LiveRange& r : l.ranges) {
        range[static_cast<std::size_t>(r.node)] = &r;
    }
    out << "{\n  \"passes\": [\"build\"" << (dce ? ", \"dce\"" : "") << (fuse ? ", \"fuse\"" : "")
        << "],\n  \"nodes\": [";
    for (const Node& n : g.nodes) {
        const std::size_t i = static_cast<std::size_t>(n.id);
        out << (n.id == 0 ? "\n" : ",\n") << "    {\"id\": " << n.id << ", \"op\": \""
            << op_name(n.op) << "\"";
        if (n.op != Op::Emit) {
            out << ", \"type\": " << json_string(type_name(n.type))
                << ", \"rate\": " << json_string(n.type.rate_text);
        }
        if (!n.name.empty()) {
            out << ", \"name\": " << json_string(n.name);
        }
        if (!n.path.empty()) {
            out << ", \"path\": " << json_string(n.path);
        }
        if (n.op == Op::Fir || n.op == Op::Channelize) {
            out << ", \"taps\": " << json_string(n.taps);
        }
        if (n.op == Op::Channelize) {
            out << ", \"channels\": " << n.n;
        }
        if (n.op == Op::Fft) {
            out << ", \"n\": " << n.n;
        }
        out << ", \"line\": " << n.pos.line << ", \"column\": " << n.pos.col;
        if (!g.live[i]) {
            out << ", \"live\": false, \"removed_by\": \"dce\"}";
            continue;
        }
        out << ", \"live\": true, \"step\": " << step[i];
        if (n.group >= 0) {
            out << ", \"group\": " << n.group;
        }
        if (const LiveRange* r = range[i]) {
            out << ", \"live_range\": [" << r->def << ", " << r->last << "], \"slot\": " << r->slot;
            if (bytes[i] >= 0) {
                out << ", \"bytes\": " << bytes[i];
            }
        }
        out << "}";
    }
    out << "\n  ],\n  \"edges\": [";
    bool first = true;
    for (const Node& n : g.nodes) {
        for (std::size_t k = 0; k < n.inputs.size(); ++k) {
            out << (first ? "\n" : ",\n") << "    {\"from\": " << n.inputs[k] << ", \"to\": " << n.id
                << ", \"input\": " << k << "}";
            first = false;
        }
    }
    out << "\n  ],\n  \"schedule\": [";
    for (std::size_t k = 0; k < s.order.size(); ++k) {
        out << (k == 0 ? "" : ", ") << s.order[k];
    }
    out << "],\n  \"slots\": " << l.slots << "\n}\n";
}

}  // namespace sigflow
