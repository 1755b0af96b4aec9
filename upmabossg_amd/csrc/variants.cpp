// Variants file parser (format in variants.h).
#include "variants.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <map>
#include <set>

#include "front.h"

namespace mbg {
namespace {

struct Tok {
    enum Kind { Ident, Number, Punct, End } kind;
    std::string text;
    int line;
};

bool ident_start(char ch) { return std::isalpha((unsigned char)ch) || ch == '_'; }
bool ident_char(char ch) { return std::isalnum((unsigned char)ch) || ch == '_' || ch == '-'; }

std::vector<Tok> lex(const std::string& s) {
    std::vector<Tok> out;
    int line = 1;
    size_t i = 0;
    while (i < s.size()) {
        const char ch = s[i];
        if (ch == '\n') {
            line++;
            i++;
        } else if (std::isspace((unsigned char)ch)) {
            i++;
        } else if (ch == '/' && i + 1 < s.size() && s[i + 1] == '/') {
            while (i < s.size() && s[i] != '\n') i++;
        } else if (ident_start(ch)) {
            size_t j = i + 1;
            while (j < s.size() && ident_char(s[j])) j++;
            out.push_back({Tok::Ident, s.substr(i, j - i), line});
            i = j;
        } else if (std::isdigit((unsigned char)ch) || (ch == '.' && i + 1 < s.size() && std::isdigit((unsigned char)s[i + 1]))) {
            size_t j = i;
            while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.')) j++;
            if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
                size_t k = j + 1;
                if (k < s.size() && (s[k] == '+' || s[k] == '-')) k++;
                if (k < s.size() && std::isdigit((unsigned char)s[k])) {
                    j = k;
                    while (j < s.size() && std::isdigit((unsigned char)s[j])) j++;
                }
            }
            out.push_back({Tok::Number, s.substr(i, j - i), line});
            i = j;
        } else if (ch == '*' && i + 1 < s.size() && s[i + 1] == '=') {
            out.push_back({Tok::Punct, "*=", line});
            i += 2;
        } else if (ch == '{' || ch == '}' || ch == ';' || ch == '=' || ch == '.') {
            out.push_back({Tok::Punct, std::string(1, ch), line});
            i++;
        } else {
            throw ParseError("line " + std::to_string(line) + ": unexpected character '" + std::string(1, ch) + "'");
        }
    }
    out.push_back({Tok::End, "", line});
    return out;
}

struct Parser {
    std::vector<Tok> t;
    size_t p = 0;
    const std::vector<std::string>& names;
    std::map<std::string, int> index;

    explicit Parser(const std::string& text, const std::vector<std::string>& n) : t(lex(text)), names(n) {
        for (size_t i = 0; i < names.size(); i++) index.emplace(names[i], (int)i);
    }

    [[noreturn]] void fail(const Tok& k, const std::string& msg) {
        throw ParseError("line " + std::to_string(k.line) + ": " + msg);
    }
    const Tok& peek() const { return t[p]; }
    const Tok& next() { return t[p++]; }
    void expect(const char* punct) {
        const Tok& k = next();
        if (k.kind != Tok::Punct || k.text != punct) fail(k, std::string("expected '") + punct + "'");
    }
    std::string ident(const char* what) {
        const Tok& k = next();
        if (k.kind != Tok::Ident) fail(k, std::string("expected ") + what);
        return k.text;
    }
    double number() {
        const Tok& k = next();
        if (k.kind != Tok::Number) fail(k, "expected a number");
        char* end = nullptr;
        const double v = std::strtod(k.text.c_str(), &end);
        if (!end || *end) fail(k, "malformed number '" + k.text + "'");
        return v;
    }
    uint64_t integer() {
        const Tok& k = next();
        bool digits = k.kind == Tok::Number && !k.text.empty();
        for (char ch : k.text) digits = digits && std::isdigit((unsigned char)ch);
        if (!digits) fail(k, "expected a non-negative integer");
        errno = 0;
        const unsigned long long v = std::strtoull(k.text.c_str(), nullptr, 10);
        if (errno == ERANGE) fail(k, "integer out of range");
        return v;
    }

    void statement(VariantSpec& v) {
        const Tok& head = peek();
        const std::string a = ident("a node name, sample_count or seed_pseudorandom");
        if (peek().kind == Tok::Punct && peek().text == "=") {
            next();
            if (a == "sample_count") {
                v.count = integer();
                if (v.count == 0) fail(head, "sample_count must be positive");
                v.has_count = true;
            } else if (a == "seed_pseudorandom") {
                v.seed = integer();
                v.has_seed = true;
            } else {
                fail(head, "unknown setting '" + a + "' (sample_count, seed_pseudorandom or NODE.attribute)");
            }
            expect(";");
            return;
        }
        const auto it = index.find(a);
        if (it == index.end()) fail(head, "unknown node '" + a + "'");
        const int i = it->second;
        expect(".");
        const Tok& at = peek();
        const std::string attr = ident("knockout, overexpress, rate_up, rate_down or istate");
        if (attr == "knockout") {
            v.istate[i] = 0.f;
            v.mult_up[i] = 0.f;
        } else if (attr == "overexpress") {
            v.istate[i] = 1.f;
            v.mult_down[i] = 0.f;
        } else if (attr == "rate_up" || attr == "rate_down") {
            expect("*=");
            const double f = number();
            if (!(f >= 0.0) || !std::isfinite(f) || (float)f == INFINITY)
                fail(at, "rate factor must be finite and non-negative");
            (attr == "rate_up" ? v.mult_up : v.mult_down)[i] *= (float)f;
        } else if (attr == "istate") {
            expect("=");
            const double q = number();
            if (!(q >= 0.0 && q <= 1.0)) fail(at, "istate must lie in [0, 1]");
            v.istate[i] = (float)q;
        } else {
            fail(at, "unknown attribute '" + attr + "'");
        }
        expect(";");
    }

    std::vector<VariantSpec> file() {
        std::vector<VariantSpec> out;
        std::set<std::string> seen;
        const size_t N = names.size();
        while (peek().kind != Tok::End) {
            const Tok& kw = next();
            if (kw.kind != Tok::Ident || kw.text != "variant") fail(kw, "expected 'variant NAME {'");
            const Tok& nm = peek();
            VariantSpec v;
            v.name = ident("a variant name");
            v.line = nm.line;
            if (!seen.insert(v.name).second) fail(nm, "duplicate variant name '" + v.name + "'");
            v.mult_up.assign(N, 1.f);
            v.mult_down.assign(N, 1.f);
            v.istate.assign(N, NAN);
            expect("{");
            while (!(peek().kind == Tok::Punct && peek().text == "}")) {
                if (peek().kind == Tok::End) fail(peek(), "unterminated block of variant '" + v.name + "'");
                statement(v);
            }
            next();
            out.push_back(std::move(v));
        }
        return out;
    }
};

}  // namespace

std::vector<VariantSpec> parse_variants(const std::string& text, const std::vector<std::string>& names) {
    Parser ps(text, names);
    return ps.file();
}

}  // namespace mbg
