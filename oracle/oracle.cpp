// CPU oracle: TEST INFRASTRUCTURE ONLY (see oracle.h for who may load it).
//
// An independent restatement of the reference algorithm (MaBoSSG) with this
// project's documented deviations.  Its own lexer/parser and expression
// interpreter are written from the reference grammar, not shared with the
// product, so a front-end mistake on either side shows up as a mismatch.
//
// Parity status of each function (what pins it):
//   lexer/parser/driver ...... golden model facts of the four reference models
//                               (node counts, names, istate, is_internal,
//                               constants) in tests/test_oracle.py
//   orc_philox4x32_10 ........ Random123 known-answer vectors and torch's Philox
//   orc_ln / orc_log2 ........ math.log over a sweep, |err| <= 5 ulp
//   node rates ............... brute force over all 2^N states (closed form
//                               of the fast path, src/generator.cpp:93-107)
//   step loop + stats ........ exact CTMC (scipy expm) on tiny networks at
//                               4 binomial sigma; the reference's test/*.res
//                               at 4 sigma (cellcycle, discrete time)
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Lexer: token set and longest-match rules of src/parser/scanner.ll:20-93.
// ---------------------------------------------------------------------------
enum Tok {
    T_EOF, T_MINUS, T_PLUS, T_STAR, T_SLASH, T_LP, T_RP, T_ASSIGN, T_QUEST, T_COLON,
    T_AND, T_OR, T_XOR, T_NOT, T_EQ, T_NE, T_LT, T_LE, T_GT, T_GE, T_LBRACE, T_RBRACE,
    T_DOT, T_SEMI, T_LBRACK, T_RBRACK, T_COMMA, T_NUMBER, T_FLOAT, T_ID, T_VAR, T_ALIAS
};

struct Token {
    Tok t;
    std::string s;
    float f = 0.f;
    int i = 0;
    int line = 1;
};

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

static bool ieq(const std::string& a, const char* b) {
    size_t n = strlen(b);
    if (a.size() != n) return false;
    for (size_t k = 0; k < n; k++)
        if (tolower((unsigned char)a[k]) != tolower((unsigned char)b[k])) return false;
    return true;
}

static std::vector<Token> lex(const std::string& src) {
    std::vector<Token> out;
    size_t p = 0, n = src.size();
    int line = 1;
    auto push = [&](Tok t, std::string s = "") {
        Token k;
        k.t = t;
        k.s = std::move(s);
        k.line = line;
        out.push_back(k);
    };
    while (p < n) {
        char c = src[p];
        if (c == ' ' || c == '\t' || c == '\r') { p++; continue; }
        if (c == '\n') { line++; p++; continue; }
        if (c == '/' && p + 1 < n && src[p + 1] == '/') {
            while (p < n && src[p] != '\n') p++;
            continue;
        }
        if (isalpha((unsigned char)c) || ((c == '$' || c == '@') && p + 1 < n && isalpha((unsigned char)src[p + 1]))) {
            size_t b = p;
            if (c == '$' || c == '@') p++;
            p++;
            while (p < n && (isalnum((unsigned char)src[p]) || src[p] == '_')) p++;
            std::string w = src.substr(b, p - b);
            if (c == '$') { push(T_VAR, w); continue; }
            if (c == '@') { push(T_ALIAS, w); continue; }
            if (ieq(w, "AND")) push(T_AND);
            else if (ieq(w, "OR")) push(T_OR);
            else if (ieq(w, "XOR")) push(T_XOR);
            else if (ieq(w, "NOT")) push(T_NOT);
            else if (ieq(w, "TRUE")) { push(T_NUMBER); out.back().i = 1; out.back().f = 1.f; }
            else if (ieq(w, "FALSE")) { push(T_NUMBER); out.back().i = 0; out.back().f = 0.f; }
            else push(T_ID, w);
            continue;
        }
        if (isdigit((unsigned char)c) || (c == '.' && p + 1 < n && isdigit((unsigned char)src[p + 1]))) {
            size_t b = p;
            bool is_float = false;
            while (p < n && isdigit((unsigned char)src[p])) p++;
            if (p < n && src[p] == '.') {
                is_float = true;
                p++;
                while (p < n && isdigit((unsigned char)src[p])) p++;
            }
            if (p < n && (src[p] == 'e' || src[p] == 'E')) {
                size_t q = p + 1;
                if (q < n && (src[q] == '+' || src[q] == '-')) q++;
                if (q < n && isdigit((unsigned char)src[q])) {
                    while (q < n && isdigit((unsigned char)src[q])) q++;
                    p = q;
                    is_float = true;
                }
            }
            std::string w = src.substr(b, p - b);
            if (is_float) {
                errno = 0;
                double d = strtod(w.c_str(), nullptr);
                if (errno == ERANGE) throw Error("float is out of range: " + w);
                push(T_FLOAT, w);
                out.back().f = (float)d;
            } else {
                errno = 0;
                long v = strtol(w.c_str(), nullptr, 10);
                if (errno == ERANGE || v < INT_MIN || v > INT_MAX) throw Error("integer is out of range: " + w);
                push(T_NUMBER, w);
                out.back().i = (int)v;
                out.back().f = (float)(int)v;
            }
            continue;
        }
        auto two = [&](char a, char b2) { return c == a && p + 1 < n && src[p + 1] == b2; };
        if (two('&', '&')) { push(T_AND); p += 2; continue; }
        if (two('|', '|')) { push(T_OR); p += 2; continue; }
        if (two('=', '=')) { push(T_EQ); p += 2; continue; }
        if (two('!', '=')) { push(T_NE); p += 2; continue; }
        if (two('<', '=')) { push(T_LE); p += 2; continue; }
        if (two('>', '=')) { push(T_GE); p += 2; continue; }
        Tok t;
        switch (c) {
            case '-': t = T_MINUS; break;
            case '+': t = T_PLUS; break;
            case '*': t = T_STAR; break;
            case '/': t = T_SLASH; break;
            case '(': t = T_LP; break;
            case ')': t = T_RP; break;
            case '=': t = T_ASSIGN; break;
            case '?': t = T_QUEST; break;
            case ':': t = T_COLON; break;
            case '&': t = T_AND; break;
            case '|': t = T_OR; break;
            case '^': t = T_XOR; break;
            case '!': t = T_NOT; break;
            case '<': t = T_LT; break;
            case '>': t = T_GT; break;
            case '{': t = T_LBRACE; break;
            case '}': t = T_RBRACE; break;
            case '.': t = T_DOT; break;
            case ';': t = T_SEMI; break;
            case '[': t = T_LBRACK; break;
            case ']': t = T_RBRACK; break;
            case ',': t = T_COMMA; break;
            default:
                throw Error("line " + std::to_string(line) + ": invalid character: " + std::string(1, c));
        }
        push(t);
        p++;
    }
    push(T_EOF);
    return out;
}

// ---------------------------------------------------------------------------
// Expression AST (src/parser/expressions.h:11-120).
// ---------------------------------------------------------------------------
enum Op { O_PLUS, O_MINUS, O_STAR, O_SLASH, O_AND, O_OR, O_XOR, O_NOT, O_LE, O_LT, O_GE, O_GT, O_EQ, O_NE };

struct Expr {
    enum Kind { LIT, ID, VAR, ALIAS, UN, BIN, TERN, PAREN } k;
    Op op = O_PLUS;
    float val = 0.f;
    std::string name;
    std::unique_ptr<Expr> a, b, c;
};
using ExprP = std::unique_ptr<Expr>;

struct Node {
    std::string name;
    std::vector<std::pair<std::string, ExprP>> attrs;
    float istate = 0.5f;  // src/parser/parse_types.cpp:3
    const Expr* attr(const std::string& n) const {
        for (auto& a : attrs)
            if (a.first == n) return a.second.get();  // first wins (parse_types.cpp:8)
        return nullptr;
    }
};

struct Driver {
    std::map<std::string, float> variables, constants;
    std::vector<Node> nodes;
    Driver() {  // src/parser/driver.cpp:7-14
        constants["discrete_time"] = 0;
        constants["sample_count"] = 1000000;
        constants["max_time"] = 10;
        constants["time_tick"] = 0.1f;
        constants["seed_pseudorandom"] = 0;
    }
    int find(const std::string& n) const {
        for (size_t i = 0; i < nodes.size(); i++)
            if (nodes[i].name == n) return (int)i;
        return -1;
    }
};

// Constant evaluation (expressions.cpp:10-120): identifiers/aliases are errors.
static float eval_const(const Expr* e, const Driver& d) {
    switch (e->k) {
        case Expr::LIT: return e->val;
        case Expr::PAREN: return eval_const(e->a.get(), d);
        case Expr::VAR: {
            auto it = d.variables.find(e->name);
            if (it == d.variables.end()) throw Error("unknown variable " + e->name);
            return it->second;
        }
        case Expr::ID: throw Error("identifier " + e->name + " in expression which needs to be evaluated");
        case Expr::ALIAS: throw Error("alias " + e->name + " in expression which needs to be evaluated");
        case Expr::UN: {
            float v = eval_const(e->a.get(), d);
            if (e->op == O_MINUS) return -v;
            if (e->op == O_NOT) return !v;
            return v;
        }
        case Expr::TERN: return eval_const(e->a.get(), d) ? eval_const(e->b.get(), d) : eval_const(e->c.get(), d);
        case Expr::BIN: {
            float l = eval_const(e->a.get(), d);
            if (e->op == O_AND) return (l && eval_const(e->b.get(), d)) ? 1.f : 0.f;
            if (e->op == O_OR) return (l || eval_const(e->b.get(), d)) ? 1.f : 0.f;
            float r = eval_const(e->b.get(), d);
            switch (e->op) {
                case O_PLUS: return l + r;
                case O_MINUS: return l - r;
                case O_STAR: return l * r;
                case O_SLASH: return l / r;
                case O_XOR: return ((l != 0) != (r != 0)) ? 1.f : 0.f;
                case O_EQ: return l == r;
                case O_NE: return l != r;
                case O_LE: return l <= r;
                case O_LT: return l < r;
                case O_GE: return l >= r;
                case O_GT: return l > r;
                default: break;
            }
        }
    }
    throw Error("bad expression");
}

// ---------------------------------------------------------------------------
// Parser: grammar and precedence of src/parser/parser.yy:74-158.
// ---------------------------------------------------------------------------
struct Parser {
    std::vector<Token> toks;
    size_t p = 0;
    Driver& d;
    Parser(std::vector<Token> t, Driver& dr) : toks(std::move(t)), d(dr) {}
    const Token& cur() const { return toks[p]; }
    [[noreturn]] void fail(const std::string& m) const {
        throw Error("line " + std::to_string(cur().line) + ": syntax error: " + m);
    }
    void expect(Tok t, const char* what) {
        if (cur().t != t) fail(std::string("expected ") + what);
        p++;
    }
    static int prec(Tok t) {
        switch (t) {
            case T_XOR: return 1;
            case T_OR: return 2;
            case T_AND: return 3;
            case T_EQ: case T_NE: return 4;
            case T_LT: case T_LE: case T_GT: case T_GE: return 5;
            case T_PLUS: case T_MINUS: return 6;
            case T_STAR: case T_SLASH: return 7;
            default: return -1;
        }
    }
    static Op binop(Tok t) {
        switch (t) {
            case T_XOR: return O_XOR;
            case T_OR: return O_OR;
            case T_AND: return O_AND;
            case T_EQ: return O_EQ;
            case T_NE: return O_NE;
            case T_LT: return O_LT;
            case T_LE: return O_LE;
            case T_GT: return O_GT;
            case T_GE: return O_GE;
            case T_PLUS: return O_PLUS;
            case T_MINUS: return O_MINUS;
            case T_STAR: return O_STAR;
            default: return O_SLASH;
        }
    }
    ExprP expr() {  // "?" is right-associative and binds loosest
        ExprP c = binary(1);
        if (cur().t == T_QUEST) {
            p++;
            ExprP m = expr();
            expect(T_COLON, "':'");
            ExprP r = expr();
            auto e = std::make_unique<Expr>();
            e->k = Expr::TERN;
            e->a = std::move(c);
            e->b = std::move(m);
            e->c = std::move(r);
            return e;
        }
        return c;
    }
    ExprP binary(int minp) {
        ExprP l = unary();
        for (;;) {
            int pr = prec(cur().t);
            if (pr < minp) return l;
            Op op = binop(cur().t);
            p++;
            ExprP r = binary(pr + 1);
            auto e = std::make_unique<Expr>();
            e->k = Expr::BIN;
            e->op = op;
            e->a = std::move(l);
            e->b = std::move(r);
            l = std::move(e);
        }
    }
    ExprP unary() {
        Tok t = cur().t;
        if (t == T_MINUS || t == T_PLUS || t == T_NOT) {
            p++;
            auto e = std::make_unique<Expr>();
            e->k = Expr::UN;
            e->op = t == T_MINUS ? O_MINUS : t == T_PLUS ? O_PLUS : O_NOT;
            e->a = unary();
            return e;
        }
        return primary();
    }
    ExprP primary() {
        auto e = std::make_unique<Expr>();
        const Token& k = cur();
        switch (k.t) {
            case T_FLOAT: case T_NUMBER: e->k = Expr::LIT; e->val = k.f; p++; return e;
            case T_ID: e->k = Expr::ID; e->name = k.s; p++; return e;
            case T_VAR: e->k = Expr::VAR; e->name = k.s; p++; return e;
            case T_ALIAS: e->k = Expr::ALIAS; e->name = k.s; p++; return e;
            case T_LP: {
                p++;
                e->k = Expr::PAREN;
                e->a = expr();
                expect(T_RP, "')'");
                return e;
            }
            default: fail("unexpected token in expression");
        }
    }
    void bnd() {
        while (cur().t != T_EOF) {
            if (cur().t != T_ID) fail("expected 'node'");
            std::string kw = cur().s;
            std::transform(kw.begin(), kw.end(), kw.begin(), ::tolower);
            if (kw != "node") fail("expected 'node' keyword");
            p++;
            if (cur().t != T_ID) fail("expected node name");
            Node nd;
            nd.name = cur().s;
            p++;
            expect(T_LBRACE, "'{'");
            do {
                if (cur().t != T_ID) fail("expected attribute name");
                std::string an = cur().s;
                p++;
                expect(T_ASSIGN, "'='");
                ExprP e = expr();
                expect(T_SEMI, "';'");
                nd.attrs.emplace_back(an, std::move(e));
            } while (cur().t != T_RBRACE);
            p++;
            // driver.cpp:45-60
            if (d.find(nd.name) >= 0) throw Error("Node " + nd.name + " already exists");
            if (!nd.attr("rate_up")) throw Error("Node " + nd.name + " does not have rate_up attribute");
            if (!nd.attr("rate_down")) throw Error("Node " + nd.name + " does not have rate_down attribute");
            d.nodes.push_back(std::move(nd));
        }
    }
    void cfg() {
        while (cur().t != T_EOF) {
            if (cur().t == T_VAR) {
                std::string n = cur().s;
                p++;
                expect(T_ASSIGN, "'='");
                ExprP e = expr();
                expect(T_SEMI, "';'");
                d.variables[n] = eval_const(e.get(), d);
            } else if (cur().t == T_ID) {
                std::string n = cur().s;
                p++;
                if (cur().t == T_DOT) {
                    p++;
                    if (cur().t != T_ID) fail("expected attribute name");
                    std::string an = cur().s;
                    p++;
                    expect(T_ASSIGN, "'='");
                    ExprP e = expr();
                    expect(T_SEMI, "';'");
                    int i = d.find(n);
                    if (i < 0) throw Error("Unknown node " + n);
                    if (an == "istate") d.nodes[i].istate = eval_const(e.get(), d);
                    else d.nodes[i].attrs.emplace_back(an, std::move(e));
                } else {
                    expect(T_ASSIGN, "'='");
                    ExprP e = expr();
                    expect(T_SEMI, "';'");
                    d.constants[n] = eval_const(e.get(), d);
                }
            } else if (cur().t == T_LBRACK) {
                // "[N].istate = a[1], b[0];" (parser.yy:102-108, driver.cpp:76-99)
                p++;
                if (cur().t != T_ID) fail("expected node name");
                std::string n = cur().s;
                p++;
                expect(T_RBRACK, "']'");
                expect(T_DOT, "'.'");
                if (cur().t != T_ID || cur().s != "istate") fail("expected 'istate' keyword");
                p++;
                expect(T_ASSIGN, "'='");
                ExprP l = expr();
                expect(T_LBRACK, "'['");
                if (cur().t != T_NUMBER) fail("expected number");
                int vl = cur().i;
                p++;
                expect(T_RBRACK, "']'");
                expect(T_COMMA, "','");
                ExprP r = expr();
                expect(T_LBRACK, "'['");
                if (cur().t != T_NUMBER) fail("expected number");
                int vr = cur().i;
                p++;
                expect(T_RBRACK, "']'");
                expect(T_SEMI, "';'");
                if (!((vl == 0 && vr == 1) || (vl == 1 && vr == 0))) fail("numbers in [] must be 0 and 1");
                int i = d.find(n);
                if (i < 0) throw Error("Unknown node " + n);
                float lp = eval_const(l.get(), d), rp = eval_const(r.get(), d);
                d.nodes[i].istate = vl == 1 ? lp / (lp + rp) : rp / (lp + rp);
            } else {
                fail("unexpected token");
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Postfix programs for state-dependent evaluation (the semantics of the code
// src/generator.cpp:84-122 emits, with logical 0/1 results for & | ! ^).
// ---------------------------------------------------------------------------
enum POp : uint8_t {
    P_CONST, P_BIT, P_NEG, P_NOT, P_ADD, P_SUB, P_MUL, P_DIV, P_AND, P_OR, P_XOR,
    P_EQ, P_NE, P_LT, P_LE, P_GT, P_GE, P_SEL
};
struct PIns {
    POp op;
    int node;
    float v;
};
using Prog = std::vector<PIns>;

static float run_prog(const Prog& pr, const uint64_t* x, float* st) {
    int sp = 0;
    for (const PIns& in : pr) {
        switch (in.op) {
            case P_CONST: st[sp++] = in.v; break;
            case P_BIT: st[sp++] = (float)((x[in.node >> 6] >> (in.node & 63)) & 1u); break;
            case P_NEG: st[sp - 1] = -st[sp - 1]; break;
            case P_NOT: st[sp - 1] = st[sp - 1] == 0.f ? 1.f : 0.f; break;
            case P_SEL: {
                float r = st[--sp], m = st[--sp], c = st[sp - 1];
                st[sp - 1] = c != 0.f ? m : r;
                break;
            }
            default: {
                float r = st[--sp], l = st[sp - 1], o;
                switch (in.op) {
                    case P_ADD: o = l + r; break;
                    case P_SUB: o = l - r; break;
                    case P_MUL: o = l * r; break;
                    case P_DIV: o = l / r; break;
                    case P_AND: o = (l != 0.f && r != 0.f) ? 1.f : 0.f; break;
                    case P_OR: o = (l != 0.f || r != 0.f) ? 1.f : 0.f; break;
                    case P_XOR: o = ((l != 0.f) != (r != 0.f)) ? 1.f : 0.f; break;
                    case P_EQ: o = l == r; break;
                    case P_NE: o = l != r; break;
                    case P_LT: o = l < r; break;
                    case P_LE: o = l <= r; break;
                    case P_GT: o = l > r; break;
                    default: o = l >= r; break;
                }
                st[sp - 1] = o;
            }
        }
    }
    return st[0];
}

struct Compiler {
    const Driver& d;
    int self;
    std::set<int>* refs;
    int depth = 0;  // nested @alias references (a self-referencing alias never ends)
    void emit(const Expr* e, Prog& pr) {
        switch (e->k) {
            case Expr::LIT: pr.push_back({P_CONST, 0, e->val}); break;
            case Expr::VAR: pr.push_back({P_CONST, 0, eval_const(e, d)}); break;
            case Expr::PAREN: emit(e->a.get(), pr); break;
            case Expr::ID: {
                int i = d.find(e->name);
                if (i < 0) throw Error("unknown node name: " + e->name);
                if (refs) refs->insert(i);
                pr.push_back({P_BIT, i, 0.f});
                break;
            }
            case Expr::ALIAS: {  // expressions.cpp:223-231
                const Expr* a = d.nodes[self].attr(e->name.substr(1));
                if (!a) throw Error("Node " + d.nodes[self].name + " does not have attribute " + e->name.substr(1));
                if (++depth > 64) throw Error("alias recursion too deep");
                emit(a, pr);
                depth--;
                break;
            }
            case Expr::UN:
                emit(e->a.get(), pr);
                if (e->op == O_MINUS) pr.push_back({P_NEG, 0, 0.f});
                else if (e->op == O_NOT) pr.push_back({P_NOT, 0, 0.f});
                break;
            case Expr::TERN:
                emit(e->a.get(), pr);
                emit(e->b.get(), pr);
                emit(e->c.get(), pr);
                pr.push_back({P_SEL, 0, 0.f});
                break;
            case Expr::BIN: {
                emit(e->a.get(), pr);
                emit(e->b.get(), pr);
                static const POp m[] = {P_ADD, P_SUB, P_MUL, P_DIV, P_AND, P_OR, P_XOR, P_NOT,
                                        P_LE, P_LT, P_GE, P_GT, P_EQ, P_NE};
                pr.push_back({m[e->op], 0, 0.f});
                break;
            }
        }
    }
};

// Fast path test of src/generator.cpp:57-82: "@logic ? U : 0" / "@logic ? 0 : D".
static bool fast_form(const Driver& d, const Expr* e, bool up, float& val) {
    if (!e || e->k != Expr::TERN) return false;
    if (e->a->k != Expr::ALIAS || e->a->name != "@logic") return false;
    float l, r;
    try {
        l = eval_const(e->b.get(), d);
        r = eval_const(e->c.get(), d);
    } catch (const Error&) {
        return false;
    }
    float zero = up ? r : l;
    val = up ? l : r;
    return zero == 0.f;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), own implementation.
// ---------------------------------------------------------------------------
static inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t o[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

// ---------------------------------------------------------------------------
// Natural log from +,-,*,/ only (fdlibm-style reduction to [sqrt(1/2), sqrt(2)),
// atanh series ln m = 2g(1 + g^2/3 + ... + g^16/17), g = (m-1)/(m+1)), so that the
// CPU and GPU produce identical bits under -ffp-contract=off.
// ---------------------------------------------------------------------------
static inline double ln_impl(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x = x * 18014398509481984.0;  // 2^54
        memcpy(&b, &x, 8);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = m + 1.0;
    double g = f / s;
    double z = g * g;
    double p = 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    double r = 2.0 * g;
    double lnm = r + r * (z * p);
    double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnm);
}
static inline double log2_impl(double x) { return ln_impl(x) * 1.4426950408889634; }
static inline double neglog_u32(uint32_t v) { return -ln_impl(((double)v + 1.0) * 2.3283064365386963e-10); }

static int ceil_log2_u64(uint64_t v) {
    int e = 0;
    while (e < 64 && (1ull << e) < v) e++;
    return e;
}
static int ceil_log2_d(double q) {
    int e = 0;
    while (ldexp(1.0, e) < q) e++;
    while (e > -1000 && ldexp(1.0, e - 1) >= q) e--;
    return e;
}

// ---------------------------------------------------------------------------
// Compiled model.
// ---------------------------------------------------------------------------
struct NodeC {
    bool fast = false;
    float up = 0.f, down = 0.f;   // fast path constants (after multipliers)
    Prog logic;                   // fast path logic
    Prog rate_up, rate_down;      // generic
    std::vector<int> refs;        // referenced nodes (for tabulation)
    bool internal = false;
    float istate = 0.5f;
    float mult_up = 1.f, mult_down = 1.f;
};

}  // namespace

struct orc_model {
    Driver d;
    std::vector<NodeC> nc;
    int N = 0, NW = 0;
    std::vector<int> readouts;
    int E = 0;
    int sigma = 0;  // entropy fixed-point exponent
    size_t max_stack = 8;

    float rate_f32(int i, const uint64_t* x, float* st) const {
        const NodeC& c = nc[i];
        bool xi = (x[i >> 6] >> (i & 63)) & 1u;
        if (c.fast) {
            bool lg = run_prog(c.logic, x, st) != 0.f;
            return xi == lg ? 0.f : (xi ? c.down : c.up);
        }
        float v = xi ? run_prog(c.rate_down, x, st) : run_prog(c.rate_up, x, st);
        return xi ? v * c.mult_down : v * c.mult_up;
    }
    uint64_t to_fixed(float v) const {
        if (v == 0.f) return 0;
        double r = nearbyint(ldexp((double)v, E));
        if (r < 1.0) r = 1.0;
        return (uint64_t)r;
    }
};

namespace {

static void check_rate(float v, const std::string& n) {
    if (!(v >= 0.f) || std::isinf(v)) throw Error("node " + n + ": rate must be finite and non-negative");
}

// Fixed-point exponent E: every nonzero rate value is an integer multiple of
// 2^-E when possible, capped so that N * max_rate * 2^E < 2^62.
static void compute_exponent(orc_model& m) {
    std::vector<float> vals;
    std::vector<std::vector<float>> node_vals(m.N);
    std::vector<float> st(m.max_stack + 4);
    for (int i = 0; i < m.N; i++) {
        const NodeC& c = m.nc[i];
        if (c.fast) {
            if (c.up != 0.f) vals.push_back(c.up);
            if (c.down != 0.f) vals.push_back(c.down);
            node_vals[i] = {c.up, c.down};
            check_rate(c.up, m.d.nodes[i].name);
            check_rate(c.down, m.d.nodes[i].name);
            continue;
        }
        std::vector<int> vars = c.refs;
        if (std::find(vars.begin(), vars.end(), i) == vars.end()) vars.push_back(i);
        if (vars.size() > 20) throw Error("node " + m.d.nodes[i].name + ": too many inputs for tabulation");
        std::vector<uint64_t> x(m.NW, 0);
        for (uint64_t a = 0; a < (1ull << vars.size()); a++) {
            std::fill(x.begin(), x.end(), 0);
            for (size_t k = 0; k < vars.size(); k++)
                if ((a >> k) & 1) x[vars[k] >> 6] |= 1ull << (vars[k] & 63);
            float v = m.rate_f32(i, x.data(), st.data());
            check_rate(v, m.d.nodes[i].name);
            if (v != 0.f) vals.push_back(v);
            node_vals[i].push_back(v);
        }
    }
    if (vals.empty()) {
        m.E = 0;
        m.sigma = 62;
        return;
    }
    int e_exact = INT_MIN;
    float vmax = 0.f;
    for (float v : vals) {
        int ex;
        double fr = frexp((double)v, &ex);
        uint32_t M = (uint32_t)ldexp(fr, 24);
        int tz = __builtin_ctz(M);
        int q = ex - 24 + tz;
        e_exact = std::max(e_exact, -q);
        vmax = std::max(vmax, v);
    }
    int ev;
    frexp((double)vmax, &ev);
    int en = ceil_log2_u64((uint64_t)m.N);
    m.E = std::min(e_exact, 62 - ev - en);
    // entropy exponent: the largest possible sum of r log2 r over the readouts stays below 2^62
    double tmax = 0.0;
    for (int j : m.readouts) {
        double mx = 0.0;
        for (float v : node_vals[j]) {
            const uint64_t r = m.to_fixed(v);
            if (!r) continue;
            const double rd = (double)r;
            const double t = rd * log2_impl(rd);
            if (t > mx) mx = t;
        }
        tmax = tmax + mx;
    }
    m.sigma = tmax < 1.0 ? 62 : 62 - ceil_log2_d(tmax);
}

static void build(orc_model& m) {
    Driver& d = m.d;
    m.N = (int)d.nodes.size();
    if (m.N == 0) throw Error("model has no nodes");
    m.NW = (m.N + 63) / 64;
    m.nc.assign(m.N, NodeC());
    m.readouts.clear();
    for (int i = 0; i < m.N; i++) {
        const Node& nd = d.nodes[i];
        NodeC& c = m.nc[i];
        c.istate = nd.istate;
        const Expr* ii = nd.attr("is_internal");  // parse_types.cpp:20-22
        c.internal = ii && eval_const(ii, d) != 0.f;
        if (!c.internal) m.readouts.push_back(i);
        float u, dn;
        std::set<int> refs;
        if (fast_form(d, nd.attr("rate_up"), true, u) && fast_form(d, nd.attr("rate_down"), false, dn)) {
            c.fast = true;
            c.up = u;
            c.down = dn;
            const Expr* lg = nd.attr("logic");
            if (!lg) throw Error("Node " + nd.name + " does not have attribute logic");
            Compiler cc{d, i, &refs};
            cc.emit(lg, c.logic);
        } else {
            Compiler cc{d, i, &refs};
            cc.emit(nd.attr("rate_up"), c.rate_up);
            cc.emit(nd.attr("rate_down"), c.rate_down);
        }
        c.refs.assign(refs.begin(), refs.end());
        m.max_stack = std::max(m.max_stack, c.logic.size() + c.rate_up.size() + c.rate_down.size() + 4);
    }
    if (m.readouts.size() > 63) throw Error("This executable supports a maximum of 63 non-internal nodes.");
    compute_exponent(m);
}

// ---------------------------------------------------------------------------
// Simulation.
// ---------------------------------------------------------------------------
// Square of one trajectory's total v over a window (its time in one projected state, or its
// TH x time), in the fixed-point units of a sum of squares: rint((v * scale)^2 * 2^F), where
// scale turns v into the trajectory's fraction of the window (or its window-average TH).  The
// sums of these squares give the standard errors of the window averages (ErrorProba, ErrorTH).
static inline uint64_t sq_units(uint64_t v, double scale, int F) {
    const double x = (double)v * scale;
    return (uint64_t)nearbyint(ldexp(x * x, F));
}

struct Tables {
    int W = 0, K = 0, FP = 0, FT = 0;
    uint64_t S = 0;
    std::vector<uint64_t> prob, th, fin;
    // sums over trajectories of the squared per-trajectory window totals: prob2 [W][S] (cell
    // maps above 20 readouts), th2 [W]; exponents FQ, FQT and the scales of sq_units
    int FQ = 0, FQT = 0;
    double sq_scale = 1.0, sq_scale_th = 1.0;
    std::vector<uint64_t> prob2, th2;
    std::vector<std::map<std::pair<uint32_t, uint64_t>, uint64_t>> cell2_maps;
    std::vector<uint64_t> cv2;
    // more than 20 readouts: per-thread maps of the non-zero cells ((window, state) -> sum) and
    // final states, merged after the run into sorted lists (window-major, states ascending)
    bool sparse = false;
    std::vector<std::map<std::pair<uint32_t, uint64_t>, uint64_t>> cell_maps;
    std::vector<std::map<uint64_t, uint64_t>> fin_maps;
    std::vector<uint32_t> cw;
    std::vector<uint64_t> cs, cv, fs, fv;

    void cell(int tix, int w, uint64_t s, uint64_t v) {
        if (!sparse) {
            __atomic_fetch_add(&prob[(uint64_t)w * S + s], v, __ATOMIC_RELAXED);
        } else if (v) {
            cell_maps[tix][{(uint32_t)w, s}] += v;
        }
    }
    void cell2(int tix, int w, uint64_t s, uint64_t q) {
        if (!sparse) {
            __atomic_fetch_add(&prob2[(uint64_t)w * S + s], q, __ATOMIC_RELAXED);
        } else if (q) {
            cell2_maps[tix][{(uint32_t)w, s}] += q;
        }
    }
    void final_state(int tix, uint64_t s) {
        if (!sparse) __atomic_fetch_add(&fin[s], 1, __ATOMIC_RELAXED);
        else fin_maps[tix][s] += 1;
    }
};

static inline void atomic_add(uint64_t* p, uint64_t v) { __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }

struct FixedSet {
    std::vector<std::vector<uint64_t>> states;
};

}  // namespace

struct orc_result {
    Tables t;
    uint64_t flips = 0;
    double seconds = 0;
    int NW = 0;
    std::vector<uint64_t> fixed_states, fixed_counts;
    std::vector<orc_summary> summaries;
};

namespace {

// Entropy of the readout rates (reference: src/generator.cpp:146-175), in FP64
// with the project's log2, over fixed-point rates:
//   TH = log2(R) - (sum_j r_j log2 r_j) / R,  0 when fewer than 2 readouts are active.
// with S = sum_j rint(r_j log2(r_j) 2^sigma) an exact integer (sigma from the largest
// possible sum, see compute_exponent), so that it can be maintained incrementally.
static uint64_t entropy_term(int sigma, uint64_t r) {
    if (!r) return 0;
    const double rd = (double)r;
    return (uint64_t)nearbyint(ldexp(rd * log2_impl(rd), sigma));
}

static double readout_entropy(const orc_model& m, const uint64_t* r) {
    uint64_t R = 0, S = 0;
    int act = 0;
    for (int j : m.readouts) {
        R += r[j];
        act += r[j] != 0;
        S += entropy_term(m.sigma, r[j]);
    }
    if (act < 2) return 0.0;
    const double Rd = (double)R;
    const double th = log2_impl(Rd) - ldexp((double)S, -m.sigma) / Rd;
    return th < 0.0 ? 0.0 : th;
}

static void simulate_range(const orc_model& m, const orc_params& p, uint64_t b, uint64_t e, Tables& T, int tix,
                           std::vector<std::vector<uint64_t>>* fixed, orc_summary* sums, uint64_t* flips_out) {
    const int N = m.N, NW = m.NW;
    const bool discrete = m.d.constants.at("discrete_time") != 0.f;
    const double maxt = (double)m.d.constants.at("max_time");
    const double tick = (double)m.d.constants.at("time_tick");
    const double inv_tick = 1.0 / tick;
    const double rscale = ldexp(1.0, -m.E);
    const float ratio_f = m.d.constants.at("max_time") / m.d.constants.at("time_tick");
    const uint64_t nsteps = ratio_f >= 2.f ? (uint64_t)std::floor(ratio_f) : 1;
    const uint32_t key[2] = {(uint32_t)p.seed, (uint32_t)(p.seed >> 32)};
    std::vector<uint64_t> x(NW), r(N);
    std::vector<float> st(m.max_stack + 4);
    uint64_t flips_total = 0;
    for (uint64_t id = b; id < e; id++) {
        // initial state: draw i of stream 0 (src/jit_kernels/simulation.cu:34-71)
        std::fill(x.begin(), x.end(), 0);
        uint32_t w4[4];
        for (int i = 0; i < N; i++) {
            if ((i & 3) == 0) {
                uint32_t c[4] = {(uint32_t)(i >> 2), 0u, (uint32_t)id, p.variant_id};
                philox(c, key, w4);
            }
            float u = (float)((w4[i & 3] >> 8) + 1u) * 5.9604644775390625e-08f;
            if (u <= m.nc[i].istate) x[i >> 6] |= 1ull << (i & 63);
        }
        double t = 0.0;
        uint64_t k = 0, nflip = 0, h = 0xcbf29ce484222325ull;
        bool fixed_pt = false;
        // this trajectory's totals over the current window: time (or steps) per projected state
        // visited, revisits merged, and TH x time; their squares go to prob2 / th2 when the
        // window ends
        std::vector<std::pair<uint64_t, uint64_t>> wcells;
        int cw = -1;
        uint64_t wth = 0;
        auto close_window = [&]() {
            if (cw < 0) return;
            for (const auto& c : wcells) T.cell2(tix, cw, c.first, sq_units(c.second, T.sq_scale, T.FQ));
            if (wth) atomic_add(&T.th2[cw], sq_units(wth, T.sq_scale_th, T.FQT));
            wcells.clear();
            wth = 0;
        };
        auto window_add = [&](int w, uint64_t s, uint64_t v, uint64_t thv) {
            if (w != cw) {
                close_window();
                cw = w;
            }
            if (v) {
                size_t j = 0;
                while (j < wcells.size() && wcells[j].first != s) j++;
                if (j == wcells.size()) wcells.push_back({s, v});
                else wcells[j].second += v;
            }
            wth += thv;
        };
        for (;;) {
            uint64_t total = 0;
            for (int i = 0; i < N; i++) {
                r[i] = m.to_fixed(m.rate_f32(i, x.data(), st.data()));
                total += r[i];
            }
            double t1, th = 0.0;
            uint32_t d1 = 0, d2 = 0;
            if (total == 0) {
                fixed_pt = true;
                t1 = maxt;
            } else {
                uint32_t c[4] = {(uint32_t)(k >> 1), 1u, (uint32_t)id, p.variant_id};
                philox(c, key, w4);
                d1 = w4[2 * (k & 1)];
                d2 = w4[2 * (k & 1) + 1];
                if (discrete) {
                    // step k+1 ends at (k+1)*tick; the last of nsteps steps ends at max_time
                    // (the rounding guard of simulation.cu:123-126 without FP32 drift)
                    t1 = (k + 1 >= nsteps) ? maxt : (double)(k + 1) * tick;
                } else {
                    double tot = (double)total * rscale;
                    t1 = t + neglog_u32(d1) / tot;
                }
                if (t1 > maxt) t1 = maxt;
                th = readout_entropy(m, r.data());
            }
            // statistics of the sojourn [t, t1) in state x (window_average_small.cu:50-136)
            uint64_t s = 0;
            for (size_t j = 0; j < m.readouts.size(); j++) s |= ((x[m.readouts[j] >> 6] >> (m.readouts[j] & 63)) & 1ull) << j;
            if (discrete) {
                int w = (int)floor(t * inv_tick + 0.5);
                if (w > T.W - 1) w = T.W - 1;
                const uint64_t thv = (uint64_t)nearbyint(ldexp(th, T.FT));
                T.cell(tix, w, s, 1);
                atomic_add(&T.th[w], thv);
                window_add(w, s, 1, thv);
            } else {
                int w = (int)floor(t * inv_tick);
                double bgn = t;
                while (t1 > bgn) {
                    double we = (double)(w + 1) * tick;
                    double en = t1 < we ? t1 : we;
                    double sl = en - bgn;
                    int ww = w > T.W - 1 ? T.W - 1 : w;
                    const uint64_t v = (uint64_t)nearbyint(ldexp(sl, T.FP));
                    const uint64_t thv = (uint64_t)nearbyint(ldexp(th * sl, T.FT));
                    T.cell(tix, ww, s, v);
                    atomic_add(&T.th[ww], thv);
                    window_add(ww, s, v, thv);
                    w++;
                    bgn = en;
                }
            }
            t = t1;
            if (t >= maxt) break;
            // the reference runs at most 100 records per launch: the 99th step of a launch is
            // recorded but not followed by a flip, and the next launch draws a fresh waiting
            // time for the same state (src/jit_kernels/simulation.cu:104-147)
            if (p.compat_batch_resume && (k + 1) % 99 == 0) {
                k++;
                continue;
            }
            // flip selection (simulation.cu:9-21) on exact fixed-point prefix sums
            uint64_t th32 = total >> 32, tl32 = total & 0xffffffffull;
            uint64_t target = (uint64_t)d2 * th32 + (((uint64_t)d2 * tl32) >> 32);
            uint64_t acc = 0;
            int sel = -1;
            for (int i = 0; i < N; i++) {
                acc += r[i];
                if (acc > target) {
                    sel = i;
                    break;
                }
            }
            x[sel >> 6] ^= 1ull << (sel & 63);
            h = (h ^ (uint64_t)sel) * 0x100000001b3ull;
            nflip++;
            k++;
        }
        close_window();
        uint64_t s = 0;
        for (size_t j = 0; j < m.readouts.size(); j++) s |= ((x[m.readouts[j] >> 6] >> (m.readouts[j] & 63)) & 1ull) << j;
        T.final_state(tix, s);
        if (fixed_pt && fixed) fixed->push_back(x);
        if (sums) {
            orc_summary& S = sums[id - b];
            S.flips = nflip;
            S.flip_hash = h;
            S.final_time = t;
            S.status = fixed_pt ? 2 : 1;
            S.pad = 0;
        }
        flips_total += nflip;
    }
    *flips_out = flips_total;
}

static void set_err(char* err, size_t n, const std::string& s) {
    if (err && n) {
        strncpy(err, s.c_str(), n - 1);
        err[n - 1] = 0;
    }
}

}  // namespace

extern "C" {

orc_model* orc_load_text(const char* bnd, const char* cfg, char* err, size_t errlen) {
    auto m = std::make_unique<orc_model>();
    try {
        Parser pb(lex(bnd), m->d);
        pb.bnd();
        Parser pc(lex(cfg), m->d);
        pc.cfg();
        build(*m);
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return nullptr;
    }
    return m.release();
}

void orc_free(orc_model* m) { delete m; }
int orc_num_nodes(const orc_model* m) { return m->N; }
const char* orc_node_name(const orc_model* m, int i) { return m->d.nodes[i].name.c_str(); }
float orc_node_istate(const orc_model* m, int i) { return m->nc[i].istate; }
int orc_node_is_internal(const orc_model* m, int i) { return m->nc[i].internal; }
int orc_node_is_fast(const orc_model* m, int i) { return m->nc[i].fast; }
double orc_constant(const orc_model* m, const char* n) {
    auto it = m->d.constants.find(n);
    return it == m->d.constants.end() ? NAN : it->second;
}
double orc_variable(const orc_model* m, const char* n) {
    auto it = m->d.variables.find(n);
    return it == m->d.variables.end() ? NAN : it->second;
}

int orc_set_variant(orc_model* m, const float* mu, const float* md, const float* is) {
    try {
        for (int i = 0; i < m->N; i++) {
            NodeC& c = m->nc[i];
            const Node& nd = m->d.nodes[i];
            float fu = mu ? mu[i] : 1.f, fd = md ? md[i] : 1.f;
            if (c.fast) {
                float u = 0, dn = 0;
                fast_form(m->d, nd.attr("rate_up"), true, u);
                fast_form(m->d, nd.attr("rate_down"), false, dn);
                c.up = u * fu;
                c.down = dn * fd;
            }
            c.mult_up = fu;
            c.mult_down = fd;
            c.istate = is ? is[i] : nd.istate;
        }
        compute_exponent(*m);
    } catch (const std::exception&) {
        return -1;
    }
    return 0;
}

int orc_rate_exponent(const orc_model* m) { return m->E; }
float orc_node_rate_f32(const orc_model* m, int i, const uint64_t* x) {
    std::vector<float> st(m->max_stack + 4);
    return m->rate_f32(i, x, st.data());
}
uint64_t orc_node_rate_fixed(const orc_model* m, int i, const uint64_t* x) {
    return m->to_fixed(orc_node_rate_f32(m, i, x));
}

orc_result* orc_run(orc_model* m, const orc_params* p, char* err, size_t errlen) {
    auto res = std::make_unique<orc_result>();
    try {
        Tables& T = res->t;
        const float maxt_f = m->d.constants.at("max_time");
        const float tick_f = m->d.constants.at("time_tick");
        if (!(tick_f > 0.f) || !(maxt_f > 0.f)) throw Error("max_time and time_tick must be positive");
        const bool discrete = m->d.constants.at("discrete_time") != 0.f;
        T.W = (int)std::ceil(maxt_f / tick_f);  // window_average_small.cpp:26 (float division)
        T.K = (int)m->readouts.size();
        T.sparse = T.K > 20;
        T.S = T.sparse ? 0 : 1ull << T.K;
        uint64_t n = p->n_trajectories ? p->n_trajectories : 1;
        if (discrete) {
            T.FP = 0;
            T.FT = 59 - ceil_log2_u64(n);
        } else {
            int eq = ceil_log2_d((double)n * (double)tick_f);
            T.FP = 60 - eq;
            T.FT = 57 - eq;
        }
        // squares: a fraction of a window is at most 1, a window-average TH below 64 (6 bits of
        // entropy at most), so n squares stay below 2^62
        T.FQ = 62 - ceil_log2_u64(n);
        T.FQT = 56 - ceil_log2_u64(n);
        T.sq_scale = discrete ? 1.0 : ldexp(1.0, -T.FP) / (double)tick_f;
        T.sq_scale_th = discrete ? ldexp(1.0, -T.FT) : ldexp(1.0, -T.FT) / (double)tick_f;
        T.prob.assign((uint64_t)T.W * T.S, 0);
        T.th.assign(T.W, 0);
        T.fin.assign(T.S, 0);
        T.prob2.assign((uint64_t)T.W * T.S, 0);
        T.th2.assign(T.W, 0);
        res->NW = m->NW;
        int nt = std::max(1, p->threads);
        uint64_t cnt = p->count;
        if ((uint64_t)nt > cnt) nt = (int)std::max<uint64_t>(1, cnt);
        std::vector<std::vector<std::vector<uint64_t>>> fx(nt);
        if (T.sparse) {
            T.cell_maps.resize(nt);
            T.cell2_maps.resize(nt);
            T.fin_maps.resize(nt);
        }
        std::vector<uint64_t> fl(nt, 0);
        if (p->record_summaries) res->summaries.resize(cnt);
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (int w = 0; w < nt; w++) {
            uint64_t b = p->first_id + cnt * w / nt, e = p->first_id + cnt * (w + 1) / nt;
            orc_summary* sp = p->record_summaries ? res->summaries.data() + (b - p->first_id) : nullptr;
            th.emplace_back(simulate_range, std::cref(*m), std::cref(*p), b, e, std::ref(T), w,
                            p->record_fixed ? &fx[w] : nullptr, sp, &fl[w]);
        }
        for (auto& t : th) t.join();
        auto t1 = std::chrono::steady_clock::now();
        res->seconds = std::chrono::duration<double>(t1 - t0).count();
        for (auto f : fl) res->flips += f;
        if (T.sparse) {
            std::map<std::pair<uint32_t, uint64_t>, uint64_t> cells;
            std::map<uint64_t, uint64_t> fins;
            for (auto& mp : T.cell_maps)
                for (auto& kv : mp) cells[kv.first] += kv.second;
            for (auto& mp : T.fin_maps)
                for (auto& kv : mp) fins[kv.first] += kv.second;
            std::map<std::pair<uint32_t, uint64_t>, uint64_t> cells2;
            for (auto& mp : T.cell2_maps)
                for (auto& kv : mp) cells2[kv.first] += kv.second;
            T.cell_maps.clear();
            T.cell2_maps.clear();
            T.fin_maps.clear();
            for (auto& kv : cells) {
                T.cw.push_back(kv.first.first);
                T.cs.push_back(kv.first.second);
                T.cv.push_back(kv.second);
                const auto q = cells2.find(kv.first);
                T.cv2.push_back(q == cells2.end() ? 0 : q->second);
            }
            for (auto& kv : fins) {
                T.fs.push_back(kv.first);
                T.fv.push_back(kv.second);
            }
        } else {
            for (uint64_t i = 0; i < T.prob.size(); i++)
                if (T.prob[i]) {
                    T.cw.push_back((uint32_t)(i / T.S));
                    T.cs.push_back(i % T.S);
                    T.cv.push_back(T.prob[i]);
                    T.cv2.push_back(T.prob2[i]);
                }
            for (uint64_t i = 0; i < T.fin.size(); i++)
                if (T.fin[i]) {
                    T.fs.push_back(i);
                    T.fv.push_back(T.fin[i]);
                }
        }
        // fixed points ordered as static_state_t::operator< (src/state.h:43-49)
        std::vector<std::vector<uint64_t>> all;
        for (auto& v : fx)
            for (auto& s : v) all.push_back(std::move(s));
        auto less = [](const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
            for (int i = (int)a.size() - 1; i >= 0; i--)
                if (a[i] != b[i]) return a[i] < b[i];
            return false;
        };
        std::sort(all.begin(), all.end(), less);
        for (size_t i = 0; i < all.size();) {
            size_t j = i;
            while (j < all.size() && all[j] == all[i]) j++;
            res->fixed_states.insert(res->fixed_states.end(), all[i].begin(), all[i].end());
            res->fixed_counts.push_back(j - i);
            i = j;
        }
    } catch (const std::exception& ex) {
        set_err(err, errlen, ex.what());
        return nullptr;
    }
    return res.release();
}

void orc_result_free(orc_result* r) { delete r; }
int orc_windows(const orc_result* r) { return r->t.W; }
int orc_readouts(const orc_result* r) { return r->t.K; }
int orc_prob_exponent(const orc_result* r) { return r->t.FP; }
int orc_th_exponent(const orc_result* r) { return r->t.FT; }
const uint64_t* orc_prob_table(const orc_result* r) { return r->t.prob.data(); }
const uint64_t* orc_th_table(const orc_result* r) { return r->t.th.data(); }
const uint64_t* orc_final_table(const orc_result* r) { return r->t.fin.data(); }
const uint64_t* orc_prob_sq_table(const orc_result* r) { return r->t.prob2.data(); }
const uint64_t* orc_th_sq_table(const orc_result* r) { return r->t.th2.data(); }
int orc_sq_exponent(const orc_result* r) { return r->t.FQ; }
int orc_th_sq_exponent(const orc_result* r) { return r->t.FQT; }
const uint64_t* orc_cell_squares(const orc_result* r) { return r->t.cv2.data(); }
uint64_t orc_cell_count(const orc_result* r) { return r->t.cv.size(); }
const uint32_t* orc_cell_windows(const orc_result* r) { return r->t.cw.data(); }
const uint64_t* orc_cell_states(const orc_result* r) { return r->t.cs.data(); }
const uint64_t* orc_cell_values(const orc_result* r) { return r->t.cv.data(); }
uint64_t orc_final_cell_count(const orc_result* r) { return r->t.fv.size(); }
const uint64_t* orc_final_cell_states(const orc_result* r) { return r->t.fs.data(); }
const uint64_t* orc_final_cell_counts(const orc_result* r) { return r->t.fv.data(); }
uint64_t orc_total_flips(const orc_result* r) { return r->flips; }
double orc_elapsed_seconds(const orc_result* r) { return r->seconds; }
uint64_t orc_fixed_count(const orc_result* r) { return r->fixed_counts.size(); }
int orc_state_words(const orc_result* r) { return r->NW; }
const uint64_t* orc_fixed_states(const orc_result* r) { return r->fixed_states.data(); }
const uint64_t* orc_fixed_counts(const orc_result* r) { return r->fixed_counts.data(); }
const orc_summary* orc_summaries(const orc_result* r) { return r->summaries.empty() ? nullptr : r->summaries.data(); }

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }
double orc_ln(double x) { return ln_impl(x); }
double orc_log2(double x) { return log2_impl(x); }
double orc_neglog_u32(uint32_t x) { return neglog_u32(x); }

}  // extern "C"
