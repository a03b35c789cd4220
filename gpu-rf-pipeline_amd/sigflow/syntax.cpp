// Source text, diagnostics, the lexer and the recursive-descent parser of SigFlow.
//
// Grammar (the language of the reference's compiler/README.md):
//   program   = { statement } ;           statement = ( binding | emit ) NEWLINE
//   binding   = IDENT "=" expr ;          emit      = "emit" IDENT "->" STRING
//   expr      = term { "|>" call } ;      (a |> f(x) is f(a, x))
//   term      = IDENT "(" [ args ] ")" | IDENT | NUMBER | STRING | "(" expr ")"
//   call      = IDENT [ "(" [ args ] ")" ]
//   args      = arg { "," arg } ;         arg = IDENT "=" literal | expr   (keywords last)
// Comments run from '#' to the end of the line. A line holding only blanks and a comment
// produces no NEWLINE token, and a last line without a line break still ends with one.

#include "sigflow.hpp"

#include <algorithm>
#include <cctype>
#include <cstdlib>

namespace sigflow {

Source::Source(std::string name, std::string text) : name_(std::move(name)), text_(std::move(text))
{
    starts_.push_back(0);
    for (std::size_t i = 0; i < text_.size(); ++i) {
        if (text_[i] == '\n') {
            starts_.push_back(i + 1);
        }
    }
}

std::string Source::line(int number) const
{
    if (number < 1 || static_cast<std::size_t>(number) > starts_.size()) {
        return "";
    }
    const std::size_t b = starts_[static_cast<std::size_t>(number) - 1];
    std::size_t e = text_.find('\n', b);
    if (e == std::string::npos) {
        e = text_.size();
    }
    if (e > b && text_[e - 1] == '\r') {
        --e;
    }
    return text_.substr(b, e - b);
}

void Diagnostics::add(bool error, const Pos& at, const std::string& message)
{
    errors_ += error ? 1 : 0;
    entries_.push_back(Entry{error, at, message});
}

void Diagnostics::print(std::ostream& out)
{
    std::stable_sort(entries_.begin(), entries_.end(), [](const Entry& a, const Entry& b) {
        return a.at.line != b.at.line ? a.at.line < b.at.line : a.at.col < b.at.col;
    });
    for (const Entry& e : entries_) {
        out << src_.name() << ":" << e.at.line << ":" << e.at.col << ": "
            << (e.error ? "error" : "warning") << ": " << e.message << "\n"
            << "  " << src_.line(e.at.line) << "\n"
            << "  " << std::string(static_cast<std::size_t>(e.at.col > 0 ? e.at.col - 1 : 0), ' ')
            << "^" << std::string(static_cast<std::size_t>(e.at.len > 1 ? e.at.len - 1 : 0), '~')
            << "\n";
    }
    entries_.clear();
}

// --- lexer -------------------------------------------------------------------------------------

const char* token_kind_name(Token::Kind kind)
{
    switch (kind) {
    case Token::Kind::Ident: return "IDENT";
    case Token::Kind::Emit: return "EMIT";
    case Token::Kind::Number: return "NUMBER";
    case Token::Kind::String: return "STRING";
    case Token::Kind::LParen: return "LPAREN";
    case Token::Kind::RParen: return "RPAREN";
    case Token::Kind::Comma: return "COMMA";
    case Token::Kind::Equals: return "EQUALS";
    case Token::Kind::Pipe: return "PIPE";
    case Token::Kind::Arrow: return "ARROW";
    case Token::Kind::Newline: return "NEWLINE";
    case Token::Kind::End: return "EOF";
    }
    return "?";
}

namespace {

bool ident_start(char c)
{
    return std::isalpha(static_cast<unsigned char>(c)) != 0 || c == '_';
}

bool ident_char(char c)
{
    return ident_start(c) || std::isdigit(static_cast<unsigned char>(c)) != 0;
}

bool digit(char c)
{
    return std::isdigit(static_cast<unsigned char>(c)) != 0;
}

}  // namespace

std::vector<Token> lex(const Source& src, Diagnostics& diag)
{
    const std::string& s = src.text();
    std::vector<Token> out;
    int line = 1;
    std::size_t line_start = 0;
    std::size_t i = 0;
    auto pos = [&](std::size_t at, std::size_t len) {
        return Pos{line, static_cast<int>(at - line_start) + 1, static_cast<int>(len)};
    };
    auto push = [&](Token::Kind k, std::size_t at, std::size_t len) {
        out.push_back(Token{k, s.substr(at, len), pos(at, len)});
    };
    auto line_open = [&] { return !out.empty() && out.back().kind != Token::Kind::Newline; };
    while (i < s.size()) {
        const char c = s[i];
        if (c == ' ' || c == '\t' || c == '\r') {
            ++i;
        } else if (c == '#') {
            while (i < s.size() && s[i] != '\n') {
                ++i;
            }
        } else if (c == '\n') {
            if (line_open()) {
                push(Token::Kind::Newline, i, 1);
            }
            ++i;
            ++line;
            line_start = i;
        } else if (ident_start(c)) {
            const std::size_t b = i;
            while (i < s.size() && ident_char(s[i])) {
                ++i;
            }
            const bool keyword = i - b == 4 && s.compare(b, 4, "emit") == 0;
            push(keyword ? Token::Kind::Emit : Token::Kind::Ident, b, i - b);
        } else if (digit(c)) {
            // digits [ "." digits ] [ ("e" | "E") [ "+" | "-" ] digits ]
            const std::size_t b = i;
            std::string missing;  // what a digit was expected after
            while (i < s.size() && digit(s[i])) {
                ++i;
            }
            if (i < s.size() && s[i] == '.') {
                ++i;
                if (i >= s.size() || !digit(s[i])) {
                    missing = ".";
                }
                while (i < s.size() && digit(s[i])) {
                    ++i;
                }
            }
            if (missing.empty() && i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
                const std::size_t e = i++;
                if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
                    ++i;
                }
                if (i >= s.size() || !digit(s[i])) {
                    missing = s.substr(e, i - e);
                }
                while (i < s.size() && digit(s[i])) {
                    ++i;
                }
            }
            if (!missing.empty()) {
                diag.error(pos(b, i - b), "malformed number '" + s.substr(b, i - b)
                                              + "': expected a digit after '" + missing + "'");
                continue;
            }
            push(Token::Kind::Number, b, i - b);
        } else if (c == '"') {
            const std::size_t b = i++;
            while (i < s.size() && s[i] != '"' && s[i] != '\n') {
                ++i;
            }
            if (i >= s.size() || s[i] != '"') {
                std::size_t e = i;
                if (e > b && s[e - 1] == '\r') {
                    --e;
                }
                diag.error(pos(b, e - b), "unterminated string");
                continue;
            }
            ++i;
            out.push_back(Token{Token::Kind::String, s.substr(b + 1, i - b - 2), pos(b, i - b)});
        } else if (c == '|' && i + 1 < s.size() && s[i + 1] == '>') {
            push(Token::Kind::Pipe, i, 2);
            i += 2;
        } else if (c == '-' && i + 1 < s.size() && s[i + 1] == '>') {
            push(Token::Kind::Arrow, i, 2);
            i += 2;
        } else if (c == '(' || c == ')' || c == ',' || c == '=') {
            push(c == '(' ? Token::Kind::LParen
                          : c == ')' ? Token::Kind::RParen
                                     : c == ',' ? Token::Kind::Comma : Token::Kind::Equals,
                 i, 1);
            ++i;
        } else {
            std::string msg = std::string("unexpected character '") + c + "'";
            if (c == '|') {
                msg += "; did you mean '|>'?";
            }
            diag.error(pos(i, 1), msg);
            ++i;
        }
    }
    if (line_open()) {
        out.push_back(Token{Token::Kind::Newline, "", pos(i, 1)});
    }
    out.push_back(Token{Token::Kind::End, "", pos(i, 1)});
    return out;
}

// --- parser ------------------------------------------------------------------------------------

namespace {

// Bounds that keep every later stage's recursion and passes small whatever the input: nesting
// of one expression (parentheses, call arguments and '|>' stages together) and operations
// (calls and emits) in a program.
constexpr int kMaxNesting = 256;
constexpr int kMaxOperations = 10000;

std::string describe(const Token& t)
{
    switch (t.kind) {
    case Token::Kind::Ident: return "identifier '" + t.text + "'";
    case Token::Kind::Number: return "number " + t.text;
    case Token::Kind::String: return "string \"" + t.text + "\"";
    case Token::Kind::Newline: return "end of line";
    case Token::Kind::End: return "end of file";
    default: return "'" + t.text + "'";
    }
}

class Parser {
public:
    Parser(const std::vector<Token>& t, Diagnostics& d) : t_(t), diag_(d) {}

    Program program()
    {
        Program prog;
        while (peek().kind != Token::Kind::End && operations_ <= kMaxOperations) {
            try {
                prog.stmts.push_back(statement());
                if (peek().kind != Token::Kind::Newline) {
                    fail_at(peek(), "expected end of line after the statement, found "
                                        + describe(peek()));
                }
                take();
            } catch (const Failed&) {
                depth_ = 0;
                skip_line();
            }
        }
        return prog;
    }

private:
    const Token& peek(std::size_t ahead = 0) const
    {
        const std::size_t k = i_ + ahead;
        return k < t_.size() ? t_[k] : t_.back();
    }
    const Token& take() { return t_[i_ + 1 < t_.size() ? i_++ : i_]; }
    [[noreturn]] void fail_at(const Token& t, const std::string& msg)
    {
        diag_.error(t.pos, msg);
        throw Failed{};
    }
    void skip_line()
    {
        while (peek().kind != Token::Kind::Newline && peek().kind != Token::Kind::End) {
            ++i_;
        }
        if (peek().kind == Token::Kind::Newline) {
            ++i_;
        }
    }

    Stmt statement()
    {
        Stmt st;
        const Token& first = peek();
        st.pos = first.pos;
        if (first.kind == Token::Kind::Emit) {
            count_operation(first);
            take();
            st.is_emit = true;
            if (peek().kind != Token::Kind::Ident) {
                fail_at(peek(), "expected the name of a binding after 'emit', found "
                                    + describe(peek()));
            }
            st.name_pos = peek().pos;
            st.name = take().text;
            if (peek().kind != Token::Kind::Arrow) {
                fail_at(peek(), "expected '->' after 'emit " + st.name + "', found "
                                    + describe(peek()));
            }
            take();
            if (peek().kind != Token::Kind::String) {
                fail_at(peek(), "expected an output path string after '->', found "
                                    + describe(peek()));
            }
            st.path_pos = peek().pos;
            st.path = take().text;
            return st;
        }
        if (first.kind != Token::Kind::Ident) {
            fail_at(first, "expected a binding ('name = ...') or an 'emit' statement, found "
                               + describe(first));
        }
        st.name_pos = first.pos;
        st.name = take().text;
        if (peek().kind != Token::Kind::Equals) {
            fail_at(peek(), "expected '=' after '" + st.name + "', found " + describe(peek()));
        }
        take();
        st.value = expr();
        return st;
    }

    ExprPtr expr()
    {
        const Nest level(*this, peek());
        ExprPtr e = term();
        int stages = 0;
        while (peek().kind == Token::Kind::Pipe) {
            take();
            if (peek().kind != Token::Kind::Ident) {
                fail_at(peek(), "expected an operation name after '|>', found " + describe(peek()));
            }
            if (depth_ + ++stages > kMaxNesting) {
                fail_at(peek(), "expression nested too deeply (more than "
                                    + std::to_string(kMaxNesting) + " levels)");
            }
            count_operation(peek());
            ExprPtr call = std::make_unique<Expr>();
            call->kind = Expr::Kind::Call;
            call->pos = peek().pos;
            call->text = take().text;
            if (peek().kind == Token::Kind::LParen) {
                args(*call);
            } else {
                call->bare = true;
            }
            ExprPtr pipe = std::make_unique<Expr>();
            pipe->kind = Expr::Kind::Pipe;
            pipe->pos = e->pos;
            pipe->input = std::move(e);
            pipe->call = std::move(call);
            e = std::move(pipe);
        }
        return e;
    }

    ExprPtr term()
    {
        const Token& t = peek();
        ExprPtr e = std::make_unique<Expr>();
        e->pos = t.pos;
        switch (t.kind) {
        case Token::Kind::Number:
            e->kind = Expr::Kind::Number;
            e->text = t.text;
            e->number = std::strtod(t.text.c_str(), nullptr);
            take();
            return e;
        case Token::Kind::String:
            e->kind = Expr::Kind::String;
            e->text = t.text;
            take();
            return e;
        case Token::Kind::LParen: {
            take();
            ExprPtr inner = expr();
            if (peek().kind != Token::Kind::RParen) {
                fail_at(peek(), "expected ')' to close the parenthesized expression, found "
                                    + describe(peek()));
            }
            take();
            return inner;
        }
        case Token::Kind::Ident:
            e->text = take().text;
            if (peek().kind == Token::Kind::LParen) {
                count_operation(t);
                e->kind = Expr::Kind::Call;
                args(*e);
            }
            return e;
        default:
            fail_at(t, "expected an expression, found " + describe(t));
        }
    }

    // "(" [ arg { "," arg } ] ")"
    void args(Expr& call)
    {
        take();  // (
        if (peek().kind == Token::Kind::RParen) {
            take();
            return;
        }
        for (;;) {
            if (peek().kind == Token::Kind::Ident && peek(1).kind == Token::Kind::Equals) {
                Kwarg k;
                k.name_pos = peek().pos;
                k.name = take().text;
                take();  // =
                const Token& v = peek();
                if (v.kind != Token::Kind::Number && v.kind != Token::Kind::String) {
                    fail_at(v, "expected a number or a string as the value of keyword argument '"
                                   + k.name + "', found " + describe(v));
                }
                k.value = term();
                call.kwargs.push_back(std::move(k));
            } else {
                if (!call.kwargs.empty()) {
                    fail_at(peek(), "positional argument after keyword argument '"
                                        + call.kwargs.back().name + "'");
                }
                call.args.push_back(expr());
            }
            if (peek().kind == Token::Kind::Comma) {
                take();
                continue;
            }
            if (peek().kind == Token::Kind::RParen) {
                take();
                return;
            }
            fail_at(peek(), "expected ',' or ')' in the arguments of '" + call.text + "', found "
                                + describe(peek()));
        }
    }

    // One more level of nesting for the rest of the enclosing scope.
    struct Nest {
        Parser& p;
        Nest(Parser& parser, const Token& at) : p(parser)
        {
            if (p.depth_ + 1 > kMaxNesting) {
                p.fail_at(at, "expression nested too deeply (more than "
                                  + std::to_string(kMaxNesting) + " levels)");
            }
            ++p.depth_;
        }
        ~Nest() { --p.depth_; }
    };

    void count_operation(const Token& at)
    {
        if (++operations_ > kMaxOperations) {
            fail_at(at, "program has more than " + std::to_string(kMaxOperations)
                            + " operations");
        }
    }

    const std::vector<Token>& t_;
    Diagnostics& diag_;
    std::size_t i_ = 0;
    int depth_ = 0;
    int operations_ = 0;
};

}  // namespace

Program parse(const std::vector<Token>& tokens, Diagnostics& diag)
{
    return Parser(tokens, diag).program();
}

}  // namespace sigflow
