// OpenQASM 3.0 reader and writer for Circuit, restating the reference's
// hand-written recursive-descent parser (src/qasm/parse.rs), its gate-name
// table (src/qasm/gates.rs) and its emitter (src/qasm/emit.rs) in C++.
//
// Accepted subset (as the reference): an optional `OPENQASM <version>;` and
// `include <string>;`, then in any order `qubit[n] name;` / `bit[n] name;`
// (the last declaration of each kind sets the register size; names are
// ignored), `gate ... { ... }` definitions (skipped), `reset` / `barrier`
// statements (skipped), stdgates calls (h x y z s t id, rx ry rz p/phase;
// cx/cnot/CX cz swap), `ctrl[(k)] @ <1q gate>`, `measure q[j] -> c[i];`,
// `c[i] = measure q[j];` and `if (c[i] == 1) <statement>`.  Angle arguments
// are expressions over literals, pi and tau with unary -, +, and binary
// + - * /.  Other tokens at statement start are skipped one at a time.
//
// Error positions and messages follow the reference token for token: an
// error reports the line and column of the next unread token at the point of
// failure.  Where the reference panics instead of returning an error (a
// string literal with no closing quote, a non-ASCII character outside a
// comment or string), this reader returns an error at that character; input
// that is not UTF-8 (which the reference's &str cannot hold) is rejected at
// 0:0.
#include "qasm.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <optional>
#include <string>
#include <vector>

namespace ev {
namespace {

[[noreturn]] void qasm_error(size_t line, size_t col, const std::string &msg)
{
    throw Error(EV_E_QASM, "OpenQASM 3 error at " + std::to_string(line) + ":" + std::to_string(col) + ": " + msg,
                line, col);
}

// ------------------------------------------------------------------ lexer
// parse.rs:9-44
enum class T {
    Ident, Int, Float, Str,
    KwOpenQasm, KwInclude, KwQubit, KwBit, KwMeasure, KwIf, KwGate, KwCtrl, KwPi, KwTau,
    Semi, LParen, RParen, LBrack, RBrack, LBrace, RBrace, Comma, Eq, EqEq, Arrow, Slash, Star, Plus, Minus, At,
    Eof
};

struct Tok {
    T kind = T::Eof;
    std::string text;  // Ident, Str
    uint64_t ival = 0; // Int
    double fval = 0;   // Float
    size_t line = 0, col = 0;
};

bool valid_utf8(const unsigned char *s, size_t n)
{
    size_t i = 0;
    while (i < n) {
        const unsigned c = s[i];
        size_t k = 0;
        uint32_t cp = 0, min = 0;
        if (c < 0x80) {
            ++i;
            continue;
        } else if ((c & 0xE0) == 0xC0) {
            k = 1, cp = c & 0x1F, min = 0x80;
        } else if ((c & 0xF0) == 0xE0) {
            k = 2, cp = c & 0x0F, min = 0x800;
        } else if ((c & 0xF8) == 0xF0) {
            k = 3, cp = c & 0x07, min = 0x10000;
        } else {
            return false;
        }
        for (size_t j = 1; j <= k; ++j) {
            if (i + j >= n || (s[i + j] & 0xC0) != 0x80) return false;
            cp = (cp << 6) | (s[i + j] & 0x3F);
        }
        if (cp < min || cp > 0x10FFFF || (cp >= 0xD800 && cp <= 0xDFFF)) return false;
        i += k + 1;
    }
    return true;
}

// Unicode White_Space (what str::trim_start removes), as (code point, length)
// of the UTF-8 sequence at p.
bool white_space_at(const unsigned char *p, size_t left, size_t *len)
{
    const unsigned c = p[0];
    if (c == ' ' || (c >= 0x09 && c <= 0x0D)) {
        *len = 1;
        return true;
    }
    if (c < 0x80) return false;
    uint32_t cp;
    size_t k;
    if ((c & 0xE0) == 0xC0) cp = c & 0x1F, k = 2;
    else if ((c & 0xF0) == 0xE0) cp = c & 0x0F, k = 3;
    else cp = c & 0x07, k = 4;
    if (k > left) return false;
    for (size_t j = 1; j < k; ++j) cp = (cp << 6) | (p[j] & 0x3F);
    const bool ws = cp == 0x85 || cp == 0xA0 || cp == 0x1680 || (cp >= 0x2000 && cp <= 0x200A) || cp == 0x2028 ||
                    cp == 0x2029 || cp == 0x202F || cp == 0x205F || cp == 0x3000;
    if (ws) *len = k;
    return ws;
}

bool is_digit(unsigned char c) { return c >= '0' && c <= '9'; }
bool is_alpha(unsigned char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }

// str::parse::<f64> on the scanner's output (parse.rs:201-222 shapes it as
// digits [. digits] [e [sign] digits]): an exponent with no digits is a
// parse error, which the lexer maps to 0.0 (parse.rs:163-169).
double rust_f64(const std::string &s)
{
    const size_t e = s.find_first_of("eE");
    if (e != std::string::npos) {
        size_t d = e + 1;
        if (d < s.size() && (s[d] == '+' || s[d] == '-')) ++d;
        if (d >= s.size()) return 0.0;
    }
    double v = 0;
    std::from_chars(s.data(), s.data() + s.size(), v);  // correctly rounded, locale-free
    return v;
}

class Lexer {
  public:
    Lexer(const char *src, size_t len) : s_(reinterpret_cast<const unsigned char *>(src)), n_(len) {}

    // parse.rs:108-199
    Tok next_token()
    {
        skip_ws_and_comments();
        Tok t;
        t.line = line_;
        t.col = col_;
        if (pos_ >= n_) return t;  // Eof
        const unsigned char ch = s_[pos_];
        auto single = [&](T k, size_t w) {
            advance(w);
            t.kind = k;
            return t;
        };
        if (ch == '"') {  // string literal
            size_t end = pos_ + 1;
            while (end < n_ && s_[end] != '"') ++end;
            if (end >= n_) qasm_error(t.line, t.col, "unterminated string literal");
            t.kind = T::Str;
            t.text.assign(reinterpret_cast<const char *>(s_ + pos_ + 1), end - pos_ - 1);
            advance(end - pos_ + 1);
            return t;
        }
        if (rem_starts("==")) return single(T::EqEq, 2);
        if (rem_starts("->")) return single(T::Arrow, 2);
        switch (ch) {
        case ';': return single(T::Semi, 1);
        case '(': return single(T::LParen, 1);
        case ')': return single(T::RParen, 1);
        case '[': return single(T::LBrack, 1);
        case ']': return single(T::RBrack, 1);
        case '{': return single(T::LBrace, 1);
        case '}': return single(T::RBrace, 1);
        case ',': return single(T::Comma, 1);
        case '=': return single(T::Eq, 1);
        case '/': return single(T::Slash, 1);
        case '*': return single(T::Star, 1);
        case '+': return single(T::Plus, 1);
        case '-': return single(T::Minus, 1);
        case '@': return single(T::At, 1);
        default: break;
        }
        if (is_digit(ch)) {  // number
            const std::string s = scan_number();
            if (s.find_first_of(".eE") != std::string::npos) {
                t.kind = T::Float;
                t.fval = rust_f64(s);
            } else {
                uint64_t v = 0;
                const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
                if (r.ec == std::errc()) {
                    t.kind = T::Int;
                    t.ival = v;
                } else {  // beyond usize: the lexer's f64 fallback
                    t.kind = T::Float;
                    t.fval = rust_f64(s);
                }
            }
            advance(s.size());
            return t;
        }
        if (is_alpha(ch) || ch == '_') {  // identifier or keyword
            size_t end = pos_;
            while (end < n_ && (is_alpha(s_[end]) || is_digit(s_[end]) || s_[end] == '_')) ++end;
            const std::string w(reinterpret_cast<const char *>(s_ + pos_), end - pos_);
            advance(end - pos_);
            static const struct {
                const char *w;
                T k;
            } kw[] = {{"OPENQASM", T::KwOpenQasm}, {"include", T::KwInclude}, {"qubit", T::KwQubit},
                      {"bit", T::KwBit},           {"measure", T::KwMeasure}, {"if", T::KwIf},
                      {"gate", T::KwGate},         {"ctrl", T::KwCtrl},       {"pi", T::KwPi},
                      {"tau", T::KwTau}};
            t.kind = T::Ident;
            for (const auto &e : kw)
                if (w == e.w) t.kind = e.k;
            if (t.kind == T::Ident) t.text = w;
            return t;
        }
        if (ch >= 0x80) qasm_error(t.line, t.col, "non-ASCII character outside a comment or string literal");
        advance(1);  // unrecognised: skipped as an identifier "?"
        t.kind = T::Ident;
        t.text = "?";
        return t;
    }

  private:
    bool rem_starts(const char *p) const
    {
        const size_t k = std::strlen(p);
        return n_ - pos_ >= k && std::memcmp(s_ + pos_, p, k) == 0;
    }

    // parse.rs:68-79: columns count characters, not bytes
    void advance(size_t k)
    {
        for (size_t i = pos_; i < pos_ + k; ++i) {
            if (s_[i] == '\n') {
                ++line_;
                col_ = 1;
            } else if ((s_[i] & 0xC0) != 0x80) {
                ++col_;
            }
        }
        pos_ += k;
    }

    // parse.rs:81-106
    void skip_ws_and_comments()
    {
        for (;;) {
            size_t w = 0, k;
            while (pos_ + w < n_ && white_space_at(s_ + pos_ + w, n_ - pos_ - w, &k)) w += k;
            if (w) {
                advance(w);
                continue;
            }
            if (rem_starts("//")) {
                size_t end = pos_;
                while (end < n_ && s_[end] != '\n') ++end;
                advance(end < n_ ? end - pos_ + 1 : end - pos_);
                continue;
            }
            if (rem_starts("/*")) {  // the closing "*/" is searched from the comment's first byte on
                size_t end = pos_ + 1;
                while (end + 1 < n_ && !(s_[end] == '*' && s_[end + 1] == '/')) ++end;
                advance(end + 1 < n_ ? end - pos_ + 2 : n_ - pos_);
                continue;
            }
            break;
        }
    }

    // parse.rs:201-222
    std::string scan_number() const
    {
        size_t i = pos_;
        while (i < n_ && is_digit(s_[i])) ++i;
        if (i < n_ && s_[i] == '.') {
            ++i;
            while (i < n_ && is_digit(s_[i])) ++i;
        }
        if (i < n_ && (s_[i] == 'e' || s_[i] == 'E')) {
            ++i;
            if (i < n_ && (s_[i] == '+' || s_[i] == '-')) ++i;
            while (i < n_ && is_digit(s_[i])) ++i;
        }
        return std::string(reinterpret_cast<const char *>(s_ + pos_), i - pos_);
    }

    const unsigned char *s_;
    size_t n_, pos_ = 0, line_ = 1, col_ = 1;
};

// ------------------------------------------------------------ gate names
// gates.rs:114-148
Mat2 lookup_gate_1q(const std::string &name, const std::vector<double> &args, bool *ok)
{
    *ok = true;
    if (args.empty()) {
        if (name == "h") return gate::h();
        if (name == "x") return gate::x();
        if (name == "y") return gate::y();
        if (name == "z") return gate::z();
        if (name == "s") return gate::s();
        if (name == "t") return gate::t();
        if (name == "id") return gate::id();
    } else if (args.size() == 1) {
        if (name == "rx") return gate::rx(args[0]);
        if (name == "ry") return gate::ry(args[0]);
        if (name == "rz") return gate::rz(args[0]);
        if (name == "p" || name == "phase") return gate::phase(args[0]);
    }
    *ok = false;
    return {};
}

bool lookup_gate_2q(const std::string &name, Mat4 *m)
{
    if (name == "cx" || name == "cnot" || name == "CX") *m = gate::cnot();
    else if (name == "cz") *m = gate::cz();
    else if (name == "swap") *m = gate::swap();
    else return false;
    return true;
}

// ----------------------------------------------------------------- parser
// parse.rs:226-808.  Statements are appended to the circuit as they are
// parsed (an If record first, then the statement it guards), in the order
// the reference pushes its ops.
class Parser {
  public:
    Parser(const char *src, size_t len) : lex_(src, len) {}

    // parse.rs:304-352
    Circuit parse_program()
    {
        uint64_t nq = 0, ncl = 0;
        if (peek_tok().kind == T::KwOpenQasm) {
            consume();
            if (peek_tok().kind == T::Int || peek_tok().kind == T::Float) consume();
            expect_semi();
        }
        if (peek_tok().kind == T::KwInclude) {
            consume();
            consume();  // the file string
            expect_semi();
        }
        for (;;) {
            const T k = peek_tok().kind;
            if (k == T::Eof) break;
            if (k == T::KwQubit) {
                nq = parse_register_decl("expected qubit register name");
            } else if (k == T::KwBit) {
                ncl = parse_register_decl("expected classical register name");
            } else if (k == T::KwGate) {  // skipped: everything up to '{', then the braced body
                consume();
                while (peek_tok().kind != T::LBrace && peek_tok().kind != T::Eof) consume();
                skip_brace_block();
            } else {
                parse_statement();
            }
        }
        out_.set_registers(nq, ncl);
        return std::move(out_);
    }

  private:
    Tok &peek_tok()
    {
        if (!peek_) peek_ = lex_.next_token();
        return *peek_;
    }

    Tok consume()
    {
        if (peek_) {
            Tok t = std::move(*peek_);
            peek_.reset();
            return t;
        }
        return lex_.next_token();
    }

    // parse.rs:251-258: the position of the next unread token
    [[noreturn]] void err(const std::string &msg)
    {
        const Tok &t = peek_tok();
        qasm_error(t.line, t.col, msg);
    }

    void expect_semi()
    {
        if (peek_tok().kind == T::Semi) consume();
        else err("expected ';'");
    }

    void expect_lbrack()
    {
        if (peek_tok().kind == T::LBrack) consume();
        else err("expected '['");
    }

    void expect_rbrack()
    {
        if (peek_tok().kind == T::RBrack) consume();
        else err("expected ']'");
    }

    uint64_t parse_int()
    {
        const Tok t = consume();
        if (t.kind != T::Int) err("expected integer");
        return t.ival;
    }

    // parse.rs:280-302
    void skip_brace_block()
    {
        if (consume().kind != T::LBrace) err("expected '{' for gate body");
        size_t depth = 1;
        for (;;) {
            const T k = consume().kind;
            if (k == T::LBrace) {
                ++depth;
            } else if (k == T::RBrace) {
                if (--depth == 0) break;
            } else if (k == T::Eof) {
                err("unexpected EOF in gate body");
            }
        }
    }

    // parse.rs:354-377
    uint64_t parse_register_decl(const char *name_msg)
    {
        consume();  // 'qubit' / 'bit'
        expect_lbrack();
        const uint64_t n = parse_int();
        expect_rbrack();
        if (consume().kind != T::Ident) err(name_msg);
        expect_semi();
        return n;
    }

    // parse.rs:409-480.  Returns false for skipped constructs.
    bool parse_statement()
    {
        const Tok &p = peek_tok();
        switch (p.kind) {
        case T::Semi: consume(); return false;
        case T::KwMeasure: parse_measure_stmt(); return true;
        case T::KwIf: parse_if(); return true;
        case T::KwCtrl: parse_ctrl_call(); return true;
        case T::Ident: {
            const std::string name = p.text;
            if (name == "reset" || name == "barrier") {
                for (;;) {
                    const T k = consume().kind;
                    if (k == T::Semi || k == T::Eof) break;
                }
                return false;
            }
            consume();
            switch (peek_tok().kind) {
            case T::LBrack: {  // c[i] = measure q[j];
                consume();
                const uint64_t idx = parse_int();
                expect_rbrack();
                if (peek_tok().kind != T::Eq) err("unexpected token after register index");
                consume();
                if (peek_tok().kind != T::KwMeasure) err("expected 'measure' after '='");
                consume();
                const uint64_t q = parse_qubit_index();
                expect_semi();
                out_.measure(q, idx);
                return true;
            }
            case T::LParen:
            case T::Ident:
            case T::Semi:
            case T::Comma: parse_gate_call_after_name(name); return true;
            default: break;
            }
            err("unexpected token in statement");
        }
        default: break;
        }
        consume();  // unknown token: skipped
        return false;
    }

    // parse.rs:483-510
    void parse_measure_stmt()
    {
        consume();  // 'measure'
        const uint64_t q = parse_qubit_index();
        if (peek_tok().kind == T::Arrow) {
            consume();
            const uint64_t into = parse_cbit_index();
            expect_semi();
            out_.measure(q, into);
            return;
        }
        if (peek_tok().kind == T::Semi) {
            consume();
            qasm_error(0, 0, "bare 'measure' without destination is not supported");
        }
        err("expected '->' or ';' after measure operand");
    }

    // parse.rs:523-569
    void parse_if()
    {
        consume();  // 'if'
        if (consume().kind != T::LParen) err("expected '(' after 'if'");
        if (consume().kind != T::Ident) err("expected classical register in if condition");
        expect_lbrack();
        const uint64_t bit = parse_int();
        expect_rbrack();
        if (consume().kind != T::EqEq) err("expected '==' in if condition");
        const Tok one = consume();
        if (one.kind != T::Int || one.ival != 1) qasm_error(0, 0, "only 'if (c[i] == 1)' is supported");
        if (consume().kind != T::RParen) err("expected ')' to close if condition");
        out_.if_bit(bit);  // guards the statement that follows (the whole parse fails without one)
        if (!parse_statement()) err("expected statement in if body");
    }

    // parse.rs:571-625
    void parse_ctrl_call()
    {
        consume();  // 'ctrl'
        uint64_t nc = 1;
        if (peek_tok().kind == T::LParen) {
            consume();
            nc = parse_int();
            if (consume().kind != T::RParen) err("expected ')'");
        }
        if (consume().kind != T::At) err("expected '@' after ctrl");
        const Tok name = consume();
        if (name.kind != T::Ident) err("expected gate name after 'ctrl @'");
        const std::vector<double> args = parse_arg_list();
        const std::vector<uint64_t> ops = parse_operand_list();
        expect_semi();
        if (ops.size() < nc + 1 || nc + 1 == 0) err("not enough operands for ctrl gate");
        bool ok;
        const Mat2 g = lookup_gate_1q(name.text, args, &ok);
        if (!ok) err("unknown single-qubit gate '" + name.text + "'");
        EV_REQUIRE(nc <= std::numeric_limits<uint32_t>::max(), EV_E_LIMIT, "too many controls");
        out_.controlled(ops.data(), uint32_t(nc), g, ops[nc]);
    }

    // parse.rs:628-673
    void parse_gate_call_after_name(const std::string &name)
    {
        const std::vector<double> args = parse_arg_list();
        const std::vector<uint64_t> ops = parse_operand_list();
        expect_semi();
        if (ops.size() == 1) {
            bool ok;
            const Mat2 g = lookup_gate_1q(name, args, &ok);
            if (!ok) err("unknown single-qubit gate '" + name + "'");
            out_.apply1(g, ops[0]);
            return;
        }
        if (ops.size() == 2 && args.empty()) {
            Mat4 g;
            if (!lookup_gate_2q(name, &g)) err("unknown two-qubit gate '" + name + "'");
            out_.apply2(g, ops[0], ops[1]);
            return;
        }
        err("unexpected gate arity");
    }

    // parse.rs:675-696
    std::vector<double> parse_arg_list()
    {
        std::vector<double> args;
        if (peek_tok().kind != T::LParen) return args;
        consume();
        for (;;) {
            const T k = peek_tok().kind;
            if (k == T::RParen) {
                consume();
                break;
            }
            if (k == T::Comma) {
                consume();
                continue;
            }
            args.push_back(parse_expr());
        }
        return args;
    }

    // parse.rs:698-714
    std::vector<uint64_t> parse_operand_list()
    {
        std::vector<uint64_t> out;
        for (;;) {
            const T k = peek_tok().kind;
            if (k == T::Ident) {
                consume();
                out.push_back(parse_index());
            } else if (k == T::Comma) {
                consume();
            } else {
                break;
            }
        }
        return out;
    }

    uint64_t parse_index()
    {
        expect_lbrack();
        const uint64_t i = parse_int();
        expect_rbrack();
        return i;
    }

    uint64_t parse_qubit_index()
    {
        if (peek_tok().kind != T::Ident) err("expected qubit register name");
        consume();
        return parse_index();
    }

    uint64_t parse_cbit_index()
    {
        if (peek_tok().kind != T::Ident) err("expected classical register name");
        consume();
        return parse_index();
    }

    // parse.rs:744-808: angle expressions
    double parse_expr() { return parse_additive(); }

    double parse_additive()
    {
        double lhs = parse_multiplicative();
        for (;;) {
            const T k = peek_tok().kind;
            if (k == T::Plus) {
                consume();
                lhs += parse_multiplicative();
            } else if (k == T::Minus) {
                consume();
                lhs -= parse_multiplicative();
            } else {
                break;
            }
        }
        return lhs;
    }

    double parse_multiplicative()
    {
        double lhs = parse_unary();
        for (;;) {
            const T k = peek_tok().kind;
            if (k == T::Star) {
                consume();
                lhs *= parse_unary();
            } else if (k == T::Slash) {
                consume();
                const double rhs = parse_unary();
                if (rhs == 0.0) err("division by zero in angle expression");
                lhs /= rhs;
            } else {
                break;
            }
        }
        return lhs;
    }

    double parse_unary()
    {
        if (peek_tok().kind == T::Minus) {
            consume();
            return -parse_primary();
        }
        if (peek_tok().kind == T::Plus) consume();
        return parse_primary();
    }

    double parse_primary()
    {
        const Tok t = consume();
        switch (t.kind) {
        case T::Float: return t.fval;
        case T::Int: return double(t.ival);
        case T::KwPi: return 3.14159265358979323846;
        case T::KwTau: return 6.28318530717958647692;
        case T::LParen: {
            const double v = parse_expr();
            if (consume().kind != T::RParen) err("expected ')'");
            return v;
        }
        default: err("expected numeric expression");
        }
    }

    Lexer lex_;
    std::optional<Tok> peek_;
    Circuit out_;
};

// --------------------------------------------------------------- emitter
bool meq(const Mat2 &a, const Mat2 &b)
{
    for (int k = 0; k < 4; ++k)
        if (!ceq(a[k], b[k])) return false;
    return true;
}

// gates.rs:11-13: tolerance check after angle extraction (norm = hypot)
bool gate1_close(const Mat2 &a, const Mat2 &b)
{
    for (int k = 0; k < 4; ++k)
        if (!(std::hypot(a[k].re - b[k].re, a[k].im - b[k].im) < 1e-12)) return false;
    return true;
}

[[noreturn]] void no_name() { qasm_error(0, 0, "gate has no OpenQASM 3 name; use a named gate constructor"); }

// Rust's `{:.17}`: fixed notation, 17 digits after the point.
std::string angle_arg(double v)
{
    char buf[512];
    std::snprintf(buf, sizeof buf, "(%.17f)", v);
    return buf;
}

// gates.rs:28-99: (name, argument string)
std::pair<const char *, std::string> gate_info_1q(const Mat2 &g)
{
    if (meq(g, gate::x())) return {"x", ""};
    if (meq(g, gate::y())) return {"y", ""};
    if (meq(g, gate::z())) return {"z", ""};
    if (meq(g, gate::h())) return {"h", ""};
    if (meq(g, gate::s())) return {"s", ""};
    if (meq(g, gate::t())) return {"t", ""};
    if (meq(g, gate::id())) return {"id", ""};
    // rx(theta): [[cos, -i sin], [-i sin, cos]] (theta/2)
    if (g[0].im == 0.0 && g[1].re == 0.0 && g[2].re == 0.0 && g[3].im == 0.0 && std::fabs(g[0].re - g[3].re) < 1e-15) {
        const double theta = 2.0 * std::atan2(-g[2].im, g[0].re);
        if (gate1_close(g, gate::rx(theta))) return {"rx", angle_arg(theta)};
    }
    // ry(theta): [[cos, -sin], [sin, cos]] (theta/2)
    if (g[0].im == 0.0 && g[1].im == 0.0 && g[2].im == 0.0 && g[3].im == 0.0 && std::fabs(g[0].re - g[3].re) < 1e-15 &&
        std::fabs(g[1].re + g[2].re) < 1e-15) {
        const double theta = 2.0 * std::atan2(g[2].re, g[0].re);
        if (gate1_close(g, gate::ry(theta))) return {"ry", angle_arg(theta)};
    }
    // diagonal: rz(theta) = diag(e^{-i theta/2}, e^{i theta/2}), p(lambda) = diag(1, e^{i lambda})
    if (ceq(g[1], c128{0, 0}) && ceq(g[2], c128{0, 0})) {
        const double theta = -2.0 * std::atan2(g[0].im, g[0].re);
        if (gate1_close(g, gate::rz(theta))) return {"rz", angle_arg(theta)};
        const double lambda = std::atan2(g[3].im, g[3].re);
        if (gate1_close(g, gate::phase(lambda))) return {"p", angle_arg(lambda)};
    }
    no_name();
}

bool m4eq(const c128 *a, const Mat4 &b)
{
    for (int k = 0; k < 16; ++k)
        if (!ceq(a[k], b[k])) return false;
    return true;
}

// gates.rs:102-113
const char *gate_info_2q(const c128 *m)
{
    if (m4eq(m, gate::cnot())) return "cx";
    if (m4eq(m, gate::cz())) return "cz";
    if (m4eq(m, gate::swap())) return "swap";
    no_name();
}

// emit.rs:40-81; returns the index after record i's op
size_t emit_op(const Circuit &c, size_t i, int depth, std::string &buf)
{
    const Op &op = c.ops()[i];
    const c128 *m = (op.kind == OpKind::Apply1 || op.kind == OpKind::Apply2 || op.kind == OpKind::Controlled)
                        ? c.mat(op)
                        : nullptr;
    switch (op.kind) {
    case OpKind::Apply1: {
        const auto [name, args] = gate_info_1q(Mat2{m[0], m[1], m[2], m[3]});
        buf += std::string(name) + args + " q[" + std::to_string(op.q0) + "];";
        return i + 1;
    }
    case OpKind::Apply2:
        buf += std::string(gate_info_2q(m)) + " q[" + std::to_string(op.q0) + "], q[" + std::to_string(op.q1) + "];";
        return i + 1;
    case OpKind::Controlled: {
        const auto [name, args] = gate_info_1q(Mat2{m[0], m[1], m[2], m[3]});
        if (op.nctrl == 1) buf += "ctrl @ ";
        else buf += "ctrl(" + std::to_string(op.nctrl) + ") @ ";
        buf += std::string(name) + args + " ";
        const uint64_t *cs = c.ctrls(op);
        for (uint32_t k = 0; k < op.nctrl; ++k) buf += "q[" + std::to_string(cs[k]) + "], ";
        buf += "q[" + std::to_string(op.q0) + "];";
        return i + 1;
    }
    case OpKind::Measure:
        buf += "c[" + std::to_string(op.q1) + "] = measure q[" + std::to_string(op.q0) + "];";
        return i + 1;
    case OpKind::If:
        if (depth > 0) qasm_error(0, 0, "nested IfClassic is not supported in OpenQASM 3 emit");
        EV_REQUIRE(i + 1 < c.size(), EV_E_ARG, "IfClassic record with no guarded op");
        buf += "if (c[" + std::to_string(op.q0) + "] == 1) ";
        return emit_op(c, i + 1, depth + 1, buf);
    }
    return i + 1;
}

}  // namespace

Circuit parse_qasm(const char *src, size_t len)
{
    if (!valid_utf8(reinterpret_cast<const unsigned char *>(src), len)) qasm_error(0, 0, "source is not valid UTF-8");
    Parser p(src, len);
    return p.parse_program();
}

std::string emit_qasm(const Circuit &c)
{
    std::string buf = "OPENQASM 3.0;\ninclude \"stdgates.inc\";\n";
    if (c.num_qubits() > 0) buf += "qubit[" + std::to_string(c.num_qubits()) + "] q;\n";
    if (c.num_classical() > 0) buf += "bit[" + std::to_string(c.num_classical()) + "] c;\n";
    buf += "\n";
    for (size_t i = 0; i < c.size();) {
        i = emit_op(c, i, 0, buf);
        buf += "\n";
    }
    return buf;
}

}  // namespace ev
