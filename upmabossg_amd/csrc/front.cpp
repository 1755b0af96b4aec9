// Hand-written scanner and Pratt parser for the .bnd/.cfg language.
//
// Tokens follow src/parser/scanner.ll:20-93 (longest match; AND/OR/XOR/NOT/
// TRUE/FALSE are case-insensitive keywords; // comments; integers become
// NUMBER, decimals and exponents FLOAT rounded to FP32).  Statements follow
// src/parser/parser.yy:74-124 and operator precedence parser.yy:126-134.
// Semantic actions follow src/parser/driver.cpp:45-99.  Extensions over the
// reference, both documented: the "[N].istate = a[1], b[0];" form (in the
// reference grammar but unreachable because its scanner lacks [ ] ,) and XOR.
#include "front.h"

#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <cstring>

namespace mbg {

int ParsedModel::index_of(const std::string& n) const {
    for (size_t i = 0; i < nodes.size(); i++)
        if (nodes[i].name == n) return (int)i;
    return -1;
}

namespace {

enum class T {
    End, Minus, Plus, Star, Slash, LParen, RParen, Assign, Question, Colon, And, Or, Xor, Not,
    Eq, Ne, Lt, Le, Gt, Ge, LBrace, RBrace, Dot, Semi, LBracket, RBracket, Comma,
    Int, Real, Ident, Var, Alias
};

struct Tk {
    T type;
    std::string text;
    double value = 0.0;
    int line = 0;
};

class Scanner {
public:
    explicit Scanner(const std::string& s) : src_(s) {}

    std::vector<Tk> run() {
        std::vector<Tk> out;
        for (;;) {
            Tk t = next();
            out.push_back(t);
            if (t.type == T::End) return out;
        }
    }

private:
    const std::string& src_;
    size_t pos_ = 0;
    int line_ = 1;

    char peek(size_t o = 0) const { return pos_ + o < src_.size() ? src_[pos_ + o] : '\0'; }

    static bool word_is(const std::string& w, const char* kw) {
        if (w.size() != strlen(kw)) return false;
        for (size_t i = 0; i < w.size(); i++)
            if (std::toupper((unsigned char)w[i]) != kw[i]) return false;
        return true;
    }

    Tk make(T t, std::string text = std::string()) {
        Tk k;
        k.type = t;
        k.text = std::move(text);
        k.line = line_;
        return k;
    }

    Tk next() {
        for (;;) {
            char c = peek();
            if (c == '\n') {
                line_++;
                pos_++;
            } else if (c == ' ' || c == '\t' || c == '\r') {
                pos_++;
            } else if (c == '/' && peek(1) == '/') {
                while (peek() && peek() != '\n') pos_++;
            } else {
                break;
            }
        }
        const char c = peek();
        if (!c) return make(T::End);
        if (std::isalpha((unsigned char)c) || ((c == '$' || c == '@') && std::isalpha((unsigned char)peek(1)))) {
            const size_t start = pos_;
            pos_ += (c == '$' || c == '@') ? 2 : 1;
            while (std::isalnum((unsigned char)peek()) || peek() == '_') pos_++;
            std::string w = src_.substr(start, pos_ - start);
            if (c == '$') return make(T::Var, w);
            if (c == '@') return make(T::Alias, w);
            if (word_is(w, "AND")) return make(T::And);
            if (word_is(w, "OR")) return make(T::Or);
            if (word_is(w, "XOR")) return make(T::Xor);
            if (word_is(w, "NOT")) return make(T::Not);
            if (word_is(w, "TRUE") || word_is(w, "FALSE")) {
                Tk k = make(T::Int, w);
                k.value = word_is(w, "TRUE") ? 1.0 : 0.0;
                return k;
            }
            return make(T::Ident, w);
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && std::isdigit((unsigned char)peek(1)))) return number();
        static const struct {
            const char* s;
            T t;
        } ops[] = {{"&&", T::And}, {"||", T::Or}, {"==", T::Eq}, {"!=", T::Ne}, {"<=", T::Le}, {">=", T::Ge},
                   {"-", T::Minus}, {"+", T::Plus}, {"*", T::Star}, {"/", T::Slash}, {"(", T::LParen},
                   {")", T::RParen}, {"=", T::Assign}, {"?", T::Question}, {":", T::Colon}, {"&", T::And},
                   {"|", T::Or}, {"^", T::Xor}, {"!", T::Not}, {"<", T::Lt}, {">", T::Gt}, {"{", T::LBrace},
                   {"}", T::RBrace}, {".", T::Dot}, {";", T::Semi}, {"[", T::LBracket}, {"]", T::RBracket},
                   {",", T::Comma}};
        for (const auto& o : ops) {
            const size_t n = strlen(o.s);
            if (src_.compare(pos_, n, o.s) == 0) {
                pos_ += n;
                return make(o.t, o.s);
            }
        }
        throw ParseError("line " + std::to_string(line_) + ": invalid character: " + std::string(1, c));
    }

    Tk number() {
        const size_t start = pos_;
        bool real = false;
        while (std::isdigit((unsigned char)peek())) pos_++;
        if (peek() == '.') {
            real = true;
            pos_++;
            while (std::isdigit((unsigned char)peek())) pos_++;
        }
        if (peek() == 'e' || peek() == 'E') {
            size_t q = 1;
            if (peek(q) == '+' || peek(q) == '-') q++;
            if (std::isdigit((unsigned char)peek(q))) {
                pos_ += q;
                while (std::isdigit((unsigned char)peek())) pos_++;
                real = true;
            }
        }
        std::string w = src_.substr(start, pos_ - start);
        errno = 0;
        if (real) {
            double d = std::strtod(w.c_str(), nullptr);
            if (errno == ERANGE) throw ParseError("line " + std::to_string(line_) + ": float is out of range: " + w);
            Tk k = make(T::Real, w);
            k.value = (double)(float)d;
            return k;
        }
        long v = std::strtol(w.c_str(), nullptr, 10);
        if (errno == ERANGE || v > INT_MAX || v < INT_MIN)
            throw ParseError("line " + std::to_string(line_) + ": integer is out of range: " + w);
        Tk k = make(T::Int, w);
        k.value = (double)(int)v;
        return k;
    }
};

// Binding powers, loosest first: ?: < XOR < OR < AND < == != < relational < + - < * / < unary.
int infix_power(T t) {
    switch (t) {
        case T::Question: return 1;
        case T::Xor: return 2;
        case T::Or: return 3;
        case T::And: return 4;
        case T::Eq: case T::Ne: return 5;
        case T::Lt: case T::Le: case T::Gt: case T::Ge: return 6;
        case T::Plus: case T::Minus: return 7;
        case T::Star: case T::Slash: return 8;
        default: return 0;
    }
}
constexpr int kUnaryPower = 9;

Ast::Oper infix_oper(T t) {
    switch (t) {
        case T::Xor: return Ast::Xor;
        case T::Or: return Ast::Or;
        case T::And: return Ast::And;
        case T::Eq: return Ast::Eq;
        case T::Ne: return Ast::Ne;
        case T::Lt: return Ast::Lt;
        case T::Le: return Ast::Le;
        case T::Gt: return Ast::Gt;
        case T::Ge: return Ast::Ge;
        case T::Plus: return Ast::Add;
        case T::Minus: return Ast::Sub;
        case T::Star: return Ast::Mul;
        default: return Ast::Div;
    }
}

class Grammar {
public:
    Grammar(std::vector<Tk> toks, ParsedModel& m) : toks_(std::move(toks)), m_(m) {}

    void bnd_program() {
        while (at().type != T::End) {
            std::string kw = take(T::Ident, "'node'").text;
            for (auto& ch : kw) ch = (char)std::tolower((unsigned char)ch);
            if (kw != "node") fail("expected 'node' keyword");
            NodeDecl nd;
            nd.name = take(T::Ident, "node name").text;
            take(T::LBrace, "'{'");
            do {
                std::string an = take(T::Ident, "attribute name").text;
                take(T::Assign, "'='");
                AstPtr e = expression(0);
                take(T::Semi, "';'");
                nd.attrs.emplace_back(std::move(an), std::move(e));
            } while (at().type != T::RBrace);
            take(T::RBrace, "'}'");
            if (m_.index_of(nd.name) >= 0) throw ParseError("Node " + nd.name + " already exists");
            if (!nd.attr("rate_up")) throw ParseError("Node " + nd.name + " does not have rate_up attribute");
            if (!nd.attr("rate_down")) throw ParseError("Node " + nd.name + " does not have rate_down attribute");
            m_.nodes.push_back(std::move(nd));
        }
    }

    void cfg_program() {
        while (at().type != T::End) {
            if (at().type == T::Var) {
                std::string name = take(T::Var, "variable").text;
                take(T::Assign, "'='");
                AstPtr e = expression(0);
                take(T::Semi, "';'");
                m_.variables[name] = evaluate(e.get(), m_);
            } else if (at().type == T::Ident) {
                std::string name = take(T::Ident, "identifier").text;
                if (at().type == T::Dot) {
                    take(T::Dot, "'.'");
                    std::string an = take(T::Ident, "attribute name").text;
                    take(T::Assign, "'='");
                    AstPtr e = expression(0);
                    take(T::Semi, "';'");
                    NodeDecl& nd = node(name);
                    if (an == "istate") nd.istate = evaluate(e.get(), m_);
                    else nd.attrs.emplace_back(std::move(an), std::move(e));
                } else {
                    take(T::Assign, "'='");
                    AstPtr e = expression(0);
                    take(T::Semi, "';'");
                    m_.constants[name] = evaluate(e.get(), m_);
                }
            } else if (at().type == T::LBracket) {
                take(T::LBracket, "'['");
                std::string name = take(T::Ident, "node name").text;
                take(T::RBracket, "']'");
                take(T::Dot, "'.'");
                if (take(T::Ident, "'istate'").text != "istate") fail("expected 'istate' keyword");
                take(T::Assign, "'='");
                AstPtr l = expression(0);
                const int vl = bracket_number();
                take(T::Comma, "','");
                AstPtr r = expression(0);
                const int vr = bracket_number();
                take(T::Semi, "';'");
                if (!((vl == 0 && vr == 1) || (vl == 1 && vr == 0))) fail("numbers in [] must be 0 and 1");
                const float lp = evaluate(l.get(), m_), rp = evaluate(r.get(), m_);
                node(name).istate = vl == 1 ? lp / (lp + rp) : rp / (lp + rp);
            } else {
                fail("unexpected token '" + at().text + "'");
            }
        }
    }

private:
    std::vector<Tk> toks_;
    size_t p_ = 0;
    int nest_ = 0;  // expression() calls in progress
    ParsedModel& m_;

    const Tk& at() const { return toks_[p_]; }
    [[noreturn]] void fail(const std::string& what) const {
        throw ParseError("line " + std::to_string(at().line) + ": syntax error, " + what);
    }
    Tk take(T t, const char* what) {
        if (at().type != t) fail(std::string("expected ") + what);
        return toks_[p_++];
    }
    int bracket_number() {
        take(T::LBracket, "'['");
        const Tk n = take(T::Int, "number");
        take(T::RBracket, "']'");
        return (int)n.value;
    }
    NodeDecl& node(const std::string& n) {
        const int i = m_.index_of(n);
        if (i < 0) throw ParseError("Unknown node " + n);
        return m_.nodes[i];
    }

    static AstPtr leaf(Ast::Kind k) {
        auto a = std::make_unique<Ast>();
        a->kind = k;
        return a;
    }

    AstPtr prefix() {
        const Tk t = toks_[p_++];
        switch (t.type) {
            case T::Int:
            case T::Real: {
                AstPtr a = leaf(Ast::Number);
                a->number = (float)t.value;
                return a;
            }
            case T::Ident: {
                AstPtr a = leaf(Ast::Node);
                a->text = t.text;
                return a;
            }
            case T::Var: {
                AstPtr a = leaf(Ast::Variable);
                a->text = t.text;
                return a;
            }
            case T::Alias: {
                AstPtr a = leaf(Ast::Alias);
                a->text = t.text;
                return a;
            }
            case T::LParen: {
                AstPtr a = leaf(Ast::Group);
                a->kids.push_back(expression(0));
                take(T::RParen, "')'");
                bounded(a);
                return a;
            }
            case T::Minus:
            case T::Plus:
            case T::Not: {
                AstPtr a = leaf(Ast::Unary);
                a->oper = t.type == T::Minus ? Ast::Sub : t.type == T::Plus ? Ast::Pos : Ast::Not;
                a->kids.push_back(expression(kUnaryPower));
                bounded(a);
                return a;
            }
            default:
                p_--;
                fail("unexpected '" + t.text + "' in expression");
        }
    }

    // Expression trees are walked recursively (evaluation, code generation, destruction), so
    // their height is bounded.
    static constexpr int kMaxHeight = 10000;  // Bison's default parser stack (YYMAXDEPTH)
    AstPtr& bounded(AstPtr& a) {
        for (const auto& k : a->kids) a->height = std::max(a->height, k->height + 1);
        if (a->height > kMaxHeight) fail("expression nested too deeply (more than " + std::to_string(kMaxHeight) + " levels)");
        return a;
    }

    // Pratt loop: binary operators are left-associative, "?:" right-associative.
    AstPtr expression(int min_power) {
        if (++nest_ > kMaxHeight) fail("expression nested too deeply (more than " + std::to_string(kMaxHeight) + " levels)");
        AstPtr e = expression_body(min_power);
        nest_--;
        return e;
    }
    AstPtr expression_body(int min_power) {
        AstPtr lhs = prefix();
        for (;;) {
            const T op = at().type;
            const int pw = infix_power(op);
            if (pw == 0 || pw <= min_power) return lhs;
            p_++;
            if (op == T::Question) {
                AstPtr a = leaf(Ast::Ternary);
                a->kids.push_back(std::move(lhs));
                a->kids.push_back(expression(0));
                take(T::Colon, "':'");
                a->kids.push_back(expression(0));
                bounded(a);
                lhs = std::move(a);
                continue;
            }
            AstPtr a = leaf(Ast::Binary);
            a->oper = infix_oper(op);
            a->kids.push_back(std::move(lhs));
            a->kids.push_back(expression(pw));
            bounded(a);
            lhs = std::move(a);
        }
    }
};

}  // namespace

float evaluate(const Ast* e, const ParsedModel& m) {
    switch (e->kind) {
        case Ast::Number: return e->number;
        case Ast::Group: return evaluate(e->kids[0].get(), m);
        case Ast::Variable: {
            auto it = m.variables.find(e->text);
            if (it == m.variables.end()) throw ParseError("unknown variable " + e->text);
            return it->second;
        }
        case Ast::Node: throw ParseError("identifier " + e->text + " in expression which needs to be evaluated");
        case Ast::Alias: throw ParseError("alias " + e->text + " in expression which needs to be evaluated");
        case Ast::Unary: {
            const float v = evaluate(e->kids[0].get(), m);
            return e->oper == Ast::Sub ? -v : e->oper == Ast::Not ? (float)!v : v;
        }
        case Ast::Ternary:
            return evaluate(e->kids[0].get(), m) != 0.f ? evaluate(e->kids[1].get(), m)
                                                          : evaluate(e->kids[2].get(), m);
        case Ast::Binary: {
            const float l = evaluate(e->kids[0].get(), m);
            if (e->oper == Ast::And) return (l != 0.f && evaluate(e->kids[1].get(), m) != 0.f) ? 1.f : 0.f;
            if (e->oper == Ast::Or) return (l != 0.f || evaluate(e->kids[1].get(), m) != 0.f) ? 1.f : 0.f;
            const float r = evaluate(e->kids[1].get(), m);
            switch (e->oper) {
                case Ast::Add: return l + r;
                case Ast::Sub: return l - r;
                case Ast::Mul: return l * r;
                case Ast::Div: return l / r;
                case Ast::Xor: return ((l != 0.f) != (r != 0.f)) ? 1.f : 0.f;
                case Ast::Eq: return l == r ? 1.f : 0.f;
                case Ast::Ne: return l != r ? 1.f : 0.f;
                case Ast::Lt: return l < r ? 1.f : 0.f;
                case Ast::Le: return l <= r ? 1.f : 0.f;
                case Ast::Gt: return l > r ? 1.f : 0.f;
                case Ast::Ge: return l >= r ? 1.f : 0.f;
                default: break;
            }
        }
    }
    throw ParseError("malformed expression");
}

void parse_model(const std::string& bnd, const std::string& cfg, ParsedModel& out) {
    Grammar gb(Scanner(bnd).run(), out);
    gb.bnd_program();
    Grammar gc(Scanner(cfg).run(), out);
    gc.cfg_program();
}

}  // namespace mbg
