#include "model.h"

#include <map>

#include <algorithm>
#include <climits>
#include <cmath>
#include <set>

#define MBG_HD static inline
#include "device/mbg_math.h"

namespace mbg {

namespace {

// The reference's fast-path test (src/generator.cpp:57-82): "@logic ? U : 0" for rate_up and
// "@logic ? 0 : D" for rate_down, with U and D constant expressions.
bool fast_form(const ParsedModel& pm, const Ast* e, bool up, float& val) {
    if (!e || e->kind != Ast::Ternary) return false;
    const Ast* c = e->kids[0].get();
    if (c->kind != Ast::Alias || c->text != "@logic") return false;
    float l, r;
    try {
        l = evaluate(e->kids[1].get(), pm);
        r = evaluate(e->kids[2].get(), pm);
    } catch (const ParseError&) {
        return false;
    }
    val = up ? l : r;
    return (up ? r : l) == 0.f;
}

struct StateEval {
    const ParsedModel& pm;
    int self;
    const std::vector<uint8_t>& bits;  // per node 0/1
    int depth = 0;

    float eval(const Ast* e) {
        switch (e->kind) {
            case Ast::Number: return e->number;
            case Ast::Variable: return evaluate(e, pm);
            case Ast::Group: return eval(e->kids[0].get());
            case Ast::Node: {
                const int i = pm.index_of(e->text);
                if (i < 0) throw ParseError("unknown node name: " + e->text);
                return bits[i] ? 1.f : 0.f;
            }
            case Ast::Alias: {
                if (++depth > 64) throw ParseError("alias recursion too deep");
                const Ast* a = pm.nodes[self].attr(e->text.substr(1));
                if (!a) throw ParseError("Node " + pm.nodes[self].name + " does not have attribute " + e->text.substr(1));
                const float v = eval(a);
                depth--;
                return v;
            }
            case Ast::Unary: {
                const float v = eval(e->kids[0].get());
                return e->oper == Ast::Sub ? -v : e->oper == Ast::Not ? (v == 0.f ? 1.f : 0.f) : v;
            }
            case Ast::Ternary: {
                const float c = eval(e->kids[0].get());
                const float m = eval(e->kids[1].get());
                const float r = eval(e->kids[2].get());
                return c != 0.f ? m : r;
            }
            case Ast::Binary: {
                const float l = eval(e->kids[0].get());
                const float r = eval(e->kids[1].get());
                switch (e->oper) {
                    case Ast::Add: return l + r;
                    case Ast::Sub: return l - r;
                    case Ast::Mul: return l * r;
                    case Ast::Div: return l / r;
                    case Ast::And: return (l != 0.f && r != 0.f) ? 1.f : 0.f;
                    case Ast::Or: return (l != 0.f || r != 0.f) ? 1.f : 0.f;
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
};

void collect_nodes(const ParsedModel& pm, int self, const Ast* e, std::set<int>& out, int depth = 0) {
    if (depth > 64) throw ParseError("alias recursion too deep");
    switch (e->kind) {
        case Ast::Node: {
            const int i = pm.index_of(e->text);
            if (i < 0) throw ParseError("unknown node name: " + e->text);
            out.insert(i);
            return;
        }
        case Ast::Alias: {
            const Ast* a = pm.nodes[self].attr(e->text.substr(1));
            if (!a) throw ParseError("Node " + pm.nodes[self].name + " does not have attribute " + e->text.substr(1));
            collect_nodes(pm, self, a, out, depth + 1);
            return;
        }
        default:
            for (const auto& k : e->kids) collect_nodes(pm, self, k.get(), out, depth);
    }
}

// Postfix program of a fast node's logic (nodes with more than kMaxInputsFast inputs),
// evaluated on a stack of bits.  An op is one u32:
//   bits 30-31 = 0  push the state bit of node (bits 0-15), negated when bit 16 is set
//              = 1  pop two, push AND (bits 0-1 = 0), OR (1) or XOR (2)
//              = 2  NOT of the top bit (bit 0 clear); SELECT (bit 0 set): pop r, m, c and
//                   push c ? m : r
//              = 3  push the constant bit 0 (bit 29 clear), or a clause (bit 29 set): push
//                   (x[w] & mask) == value, negated when bit 28 is set, with w, mask (lo, hi)
//                   and value (lo, hi) in the 5 words after the op
// Operands of an AND (OR) chain that are literals are grouped by state word into clauses:
// x[w] & mask == value tests every literal of that word at once (for an OR, the negated
// clause of the negated literals), so a chain of 60 literals over 2 words costs 2 clauses.
// The rate only depends on whether the logic is non-zero (generator.cpp:93-107), and &, |,
// ^, ! and ?: only test their operands for zero, so numbers become constant bits and unary
// minus is the identity.  Arithmetic and comparisons are refused for such nodes.
struct ProgCompiler {
    const ParsedModel& pm;
    int self;
    std::vector<uint32_t>& out;
    int depth = 0, max_depth = 0, alias_depth = 0;

    void push_op(uint32_t op, int delta) {
        out.push_back(op);
        depth += delta;
        max_depth = std::max(max_depth, depth);
    }
    [[noreturn]] void refuse() {
        throw ParseError("node " + pm.nodes[self].name + " reads more than " + std::to_string(kMaxInputsFast) +
                         " nodes: its logic may then only use & | ^ ! ?: ( ), nodes, @aliases and constants");
    }
    // e through parentheses, unary plus / minus and @aliases
    const Ast* strip(const Ast* e) {
        for (int hops = 0;; hops++) {
            if (hops > 64) throw ParseError("alias recursion too deep");
            if (e->kind == Ast::Group || (e->kind == Ast::Unary && (e->oper == Ast::Sub || e->oper == Ast::Pos))) {
                e = e->kids[0].get();
            } else if (e->kind == Ast::Alias) {
                const Ast* a = pm.nodes[self].attr(e->text.substr(1));
                if (!a) throw ParseError("Node " + pm.nodes[self].name + " does not have attribute " + e->text.substr(1));
                e = a;
            } else {
                return e;
            }
        }
    }
    // [!]...[!]node: its index and whether it is taken positively
    bool literal(const Ast* e, int& node, bool& positive) {
        positive = true;
        for (int hops = 0; hops < 256; hops++) {
            e = strip(e);
            if (e->kind == Ast::Unary && e->oper == Ast::Not) {
                positive = !positive;
                e = e->kids[0].get();
                continue;
            }
            if (e->kind != Ast::Node) return false;
            node = pm.index_of(e->text);
            if (node < 0) throw ParseError("unknown node name: " + e->text);
            return true;
        }
        return false;
    }
    void flatten(const Ast* e, Ast::Oper op, std::vector<const Ast*>& out) {
        const Ast* s = strip(e);
        if (s->kind == Ast::Binary && s->oper == op) {
            flatten(s->kids[0].get(), op, out);
            flatten(s->kids[1].get(), op, out);
        } else {
            out.push_back(e);
        }
    }
    void push_clause(uint32_t w, uint64_t mask, uint64_t value, bool neg) {
        push_op((3u << 30) | (1u << 29) | (neg ? 1u << 28 : 0u), 1);
        out.push_back(w);
        out.push_back((uint32_t)mask);
        out.push_back((uint32_t)(mask >> 32));
        out.push_back((uint32_t)value);
        out.push_back((uint32_t)(value >> 32));
    }
    // AND / OR chain: literals grouped into one clause per state word, other operands
    // compiled on their own, combined left to right (stack depth stays at 2 above theirs)
    void emit_chain(const Ast* e, Ast::Oper op, bool negate) {
        std::vector<const Ast*> ops;
        flatten(e, op, ops);
        const bool is_or = op == Ast::Or;
        std::map<uint32_t, std::pair<uint64_t, uint64_t>> words;  // word -> (mask, required value)
        std::vector<const Ast*> rest;
        for (const Ast* o : ops) {
            int j;
            bool pos;
            if (!literal(o, j, pos)) {
                rest.push_back(o);
                continue;
            }
            // AND: every literal true (x_j == pos); OR: not every literal false (x_j == !pos)
            const bool want = is_or ? !pos : pos;
            const uint64_t bit = 1ull << (j & 63);
            auto& mv = words[(uint32_t)j >> 6];
            if (mv.first & bit) {
                if (((mv.second & bit) != 0) != want) {  // A & !A is false, A | !A is true
                    push_op((3u << 30) | ((is_or != negate) ? 1u : 0u), 1);
                    return;
                }
            } else {
                mv.first |= bit;
                if (want) mv.second |= bit;
            }
        }
        const uint32_t code = is_or ? 1u : 0u;
        bool first = true;
        for (const auto& kv : words) {
            push_clause(kv.first, kv.second.first, kv.second.second, is_or);
            if (!first) push_op((1u << 30) | code, -1);
            first = false;
        }
        for (const Ast* r : rest) {
            emit(r);
            if (!first) push_op((1u << 30) | code, -1);
            first = false;
        }
        if (negate) push_op(2u << 30, 0);
    }
    void emit(const Ast* e, bool negate = false) {
        switch (e->kind) {
            case Ast::Number:
                push_op((3u << 30) | ((e->number != 0.f) != negate ? 1u : 0u), 1);
                return;
            case Ast::Variable:
                push_op((3u << 30) | ((evaluate(e, pm) != 0.f) != negate ? 1u : 0u), 1);
                return;
            case Ast::Group:
                emit(e->kids[0].get(), negate);
                return;
            case Ast::Node: {
                const int i = pm.index_of(e->text);
                if (i < 0) throw ParseError("unknown node name: " + e->text);
                push_op((uint32_t)i | (negate ? 1u << 16 : 0u), 1);
                return;
            }
            case Ast::Alias: {
                if (++alias_depth > 64) throw ParseError("alias recursion too deep");
                const Ast* a = pm.nodes[self].attr(e->text.substr(1));
                if (!a) throw ParseError("Node " + pm.nodes[self].name + " does not have attribute " + e->text.substr(1));
                emit(a, negate);
                alias_depth--;
                return;
            }
            case Ast::Unary:
                if (e->oper == Ast::Not) {
                    emit(e->kids[0].get(), !negate);
                    return;
                }
                if (e->oper == Ast::Sub || e->oper == Ast::Add || e->oper == Ast::Pos) {
                    emit(e->kids[0].get(), negate);
                    return;
                }
                refuse();
            case Ast::Ternary:
                emit(e->kids[0].get());
                emit(e->kids[1].get(), negate);
                emit(e->kids[2].get(), negate);
                push_op((2u << 30) | 1u, -2);
                return;
            case Ast::Binary: {
                if (e->oper == Ast::And || e->oper == Ast::Or) {
                    emit_chain(e, e->oper, negate);
                    return;
                }
                uint32_t code;
                if (e->oper == Ast::Xor) code = 2;
                else refuse();
                emit(e->kids[0].get());
                emit(e->kids[1].get());
                push_op((1u << 30) | code, -1);
                if (negate) push_op(2u << 30, 0);
                return;
            }
        }
        refuse();
    }
};

void check_rate(float v, const std::string& node) {
    if (!(v >= 0.f) || std::isinf(v))
        throw ParseError("node " + node + ": rates must be finite and non-negative");
}

}  // namespace

namespace {

// Host evaluation of a logic program (the op format above), as the kernel runs it.
uint32_t run_program(const CompiledModel& cm, uint32_t off, const uint64_t* x) {
    const uint32_t end = off + 1 + cm.prog[off];
    uint64_t st = 0;
    for (uint32_t p = off + 1; p < end;) {
        const uint32_t op = cm.prog[p];
        const uint32_t kind = op >> 30;
        const uint64_t a = st & 1u, b = (st >> 1) & 1u, c = (st >> 2) & 1u;
        if (kind == 3 && ((op >> 29) & 1u)) {
            const uint64_t m = (uint64_t)cm.prog[p + 2] | (uint64_t)cm.prog[p + 3] << 32;
            const uint64_t v = (uint64_t)cm.prog[p + 4] | (uint64_t)cm.prog[p + 5] << 32;
            st = (st << 1) | (uint64_t)(((x[cm.prog[p + 1]] & m) == v) != (((op >> 28) & 1u) != 0));
            p += 6;
            continue;
        }
        if (kind == 0) {
            const uint32_t j = op & 0xffffu;
            st = (st << 1) | (((x[j >> 6] >> (j & 63)) & 1u) ^ ((op >> 16) & 1u));
        } else if (kind == 1) {
            const uint32_t f = op & 3u;
            st = ((st >> 2) << 1) | (f == 0 ? (a & b) : f == 1 ? (a | b) : (a ^ b));
        } else if (kind == 2) {
            st = (op & 1u) ? (((st >> 3) << 1) | (c ? b : a)) : (st ^ 1u);
        } else {
            st = (st << 1) | (op & 1u);
        }
        p++;
    }
    return (uint32_t)(st & 1u);
}

}  // namespace

float node_rate(const CompiledModel& cm, int i, const uint64_t* x) {
    const NodeRec& r = cm.nodes[i];
    const uint32_t xi = (uint32_t)(x[i >> 6] >> (i & 63)) & 1u;
    if (r.in_cnt == kProgInputs) {
        const uint32_t logic = run_program(cm, r.tab_off, x);
        return xi == logic ? 0.f : (xi ? cm.down[i] : cm.up[i]);
    }
    uint32_t key = xi;
    for (int m = 0; m < r.in_cnt; m++) {
        const uint32_t j = cm.inputs[r.in_off + m];
        key |= (uint32_t)((x[j >> 6] >> (j & 63)) & 1u) << (m + 1);
    }
    if (cm.generic[i]) return cm.gtab[r.tab_off + key];
    const uint64_t bp = (uint64_t)r.tab_off + key;
    const uint32_t logic = (cm.lut[bp >> 5] >> (bp & 31)) & 1u;
    return xi == logic ? 0.f : (xi ? cm.down[i] : cm.up[i]);
}

uint64_t entropy_term(int sigma, uint64_t r) {
    if (!r) return 0;
    const double rd = (double)r;
    return (uint64_t)std::nearbyint(std::ldexp(rd * mbg_log2(rd), sigma));
}

int ceil_log2_u64(uint64_t v) {
    int e = 0;
    while (e < 64 && (1ull << e) < v) e++;
    return e;
}

int ceil_log2_real(double q) {
    int e = 0;
    while (std::ldexp(1.0, e) < q) e++;
    while (e > -1000 && std::ldexp(1.0, e - 1) >= q) e--;
    return e;
}

CompiledModel compile_model(const ParsedModel& pm) {
    CompiledModel cm;
    cm.N = (int)pm.nodes.size();
    if (cm.N == 0) throw ParseError("model has no nodes");
    if (cm.N > 65535) throw ParseError("too many nodes (limit 65535)");
    cm.NW = (cm.N + 63) / 64;
    cm.discrete = pm.constant("discrete_time") != 0.f;
    cm.max_time = pm.constant("max_time");
    cm.time_tick = pm.constant("time_tick");
    cm.sample_count = pm.constant("sample_count");
    cm.seed = pm.constant("seed_pseudorandom");
    cm.nodes.resize(cm.N);
    cm.generic.assign(cm.N, 0);
    cm.internal.assign(cm.N, 0);
    cm.up.assign(cm.N, 0.f);
    cm.down.assign(cm.N, 0.f);
    cm.gen_mask.assign(cm.NW, 0);
    for (int i = 0; i < cm.N; i++) {
        const NodeDecl& nd = pm.nodes[i];
        cm.names.push_back(nd.name);
        cm.istate.push_back(nd.istate);
        const Ast* ii = nd.attr("is_internal");  // src/parser/parse_types.cpp:20-22
        cm.internal[i] = ii && evaluate(ii, pm) != 0.f;
        if (!cm.internal[i]) cm.readouts.push_back(i);
    }
    cm.K = (int)cm.readouts.size();
    for (int j = 0; j < cm.K; j++) cm.nodes[cm.readouts[j]].flags = (uint8_t)(j + 1);

    std::vector<std::vector<uint32_t>> dep_lists(cm.N);
    std::vector<uint8_t> bits(cm.N, 0);
    uint64_t lut_bits = 0;
    for (int i = 0; i < cm.N; i++) {
        const NodeDecl& nd = pm.nodes[i];
        NodeRec& rec = cm.nodes[i];
        float u = 0.f, d = 0.f;
        const bool fast = fast_form(pm, nd.attr("rate_up"), true, u) && fast_form(pm, nd.attr("rate_down"), false, d);
        std::set<int> refs;
        const Ast* logic = nd.attr("logic");
        if (fast) {
            if (!logic) throw ParseError("Node " + nd.name + " does not have attribute logic");
            collect_nodes(pm, i, logic, refs);
        } else {
            collect_nodes(pm, i, nd.attr("rate_up"), refs);
            collect_nodes(pm, i, nd.attr("rate_down"), refs);
        }
        const bool self_ref = refs.erase(i) > 0;
        const int cnt = (int)refs.size();
        const bool prog = fast && cnt > kMaxInputsFast;  // logic program instead of a truth table
        if (!fast && cnt > kMaxInputsGeneric)
            throw ParseError("node " + nd.name + " reads " + std::to_string(cnt) + " nodes (limit " +
                             std::to_string(kMaxInputsGeneric) + " for rates other than @logic ? U : 0)");
        if (!prog) cm.max_inputs = std::max(cm.max_inputs, cnt);
        rec.in_off = (uint32_t)cm.inputs.size();
        rec.in_cnt = prog ? kProgInputs : (uint8_t)cnt;
        std::vector<int> in(refs.begin(), refs.end());
        for (int m = 0; m < cnt; m++) {
            cm.inputs.push_back((uint32_t)in[m]);
            // key bit position of the input (a program node has no key: any non-zero position)
            dep_lists[in[m]].push_back((uint32_t)i | ((uint32_t)(prog ? 1 : m + 1) << 24));
        }
        // A fast node whose logic does not read itself has rate 0 right after it flips
        // (it was active: x_i != logic; flipping makes them equal), so the kernel updates it
        // without a table lookup; every other node re-evaluates itself as a dependent.
        if (!fast || self_ref) {
            dep_lists[i].push_back((uint32_t)i);
            rec.flags |= 0x40;
        }
        StateEval ev{pm, i, bits};
        if (prog) {
            check_rate(u, nd.name);
            check_rate(d, nd.name);
            cm.up[i] = u;
            cm.down[i] = d;
            cm.any_prog = true;
            rec.tab_off = (uint32_t)cm.prog.size();
            cm.prog.push_back(0u);  // length, filled below
            ProgCompiler pc{pm, i, cm.prog};
            pc.emit(logic);
            if (pc.max_depth > kMaxProgDepth)
                throw ParseError("node " + nd.name + ": logic nests too deeply (bit stack above " +
                                 std::to_string(kMaxProgDepth) + ")");
            cm.prog[rec.tab_off] = (uint32_t)(cm.prog.size() - rec.tab_off - 1);
            continue;
        }
        const uint32_t nkeys = 1u << (cnt + 1);
        if (fast) {
            check_rate(u, nd.name);
            check_rate(d, nd.name);
            cm.up[i] = u;
            cm.down[i] = d;
            rec.tab_off = (uint32_t)lut_bits;
            cm.lut.resize((lut_bits + nkeys + 31) / 32, 0u);
            for (uint32_t key = 0; key < nkeys; key++) {
                bits[i] = key & 1u;
                for (int m = 0; m < cnt; m++) bits[in[m]] = (key >> (m + 1)) & 1u;
                if (ev.eval(logic) != 0.f) cm.lut[(lut_bits + key) >> 5] |= 1u << ((lut_bits + key) & 31);
            }
            lut_bits += nkeys;
        } else {
            cm.generic[i] = 1;
            cm.any_generic = true;
            rec.flags |= 0x80;
            cm.gen_mask[i >> 6] |= 1ull << (i & 63);
            rec.tab_off = (uint32_t)cm.gtab.size();
            const Ast* ru = nd.attr("rate_up");
            const Ast* rd = nd.attr("rate_down");
            for (uint32_t key = 0; key < nkeys; key++) {
                bits[i] = key & 1u;
                for (int m = 0; m < cnt; m++) bits[in[m]] = (key >> (m + 1)) & 1u;
                const float v = (key & 1u) ? ev.eval(rd) : ev.eval(ru);
                check_rate(v, nd.name);
                cm.gtab.push_back(v);
            }
        }
        bits[i] = 0;
        for (int m = 0; m < cnt; m++) bits[in[m]] = 0;
    }
    if (cm.any_prog) cm.prog.resize(cm.prog.size() + 8, 0u);  // the evaluator reads 5 words past an op
    for (int i = 0; i < cm.N; i++) {
        if (dep_lists[i].size() > 0xffff) throw ParseError("node " + cm.names[i] + " has too many dependents");
        cm.nodes[i].dep_off = (uint32_t)cm.deps.size();
        cm.nodes[i].dep_cnt = (uint16_t)dep_lists[i].size();
        cm.deps.insert(cm.deps.end(), dep_lists[i].begin(), dep_lists[i].end());
    }
    return cm;
}

RateTables make_rate_tables(const CompiledModel& cm, const float* mult_up, const float* mult_down,
                            const float* istate) {
    RateTables t;
    std::vector<float> up(cm.N), down(cm.N), gt = cm.gtab;
    for (int i = 0; i < cm.N; i++) {
        const float fu = mult_up ? mult_up[i] : 1.f, fd = mult_down ? mult_down[i] : 1.f;
        if (!(fu >= 0.f) || !(fd >= 0.f) || std::isinf(fu) || std::isinf(fd))
            throw ParseError("rate multipliers must be finite and non-negative");
        if (cm.generic[i]) {
            const NodeRec& r = cm.nodes[i];
            const uint32_t nkeys = 1u << (r.in_cnt + 1);
            for (uint32_t key = 0; key < nkeys; key++) gt[r.tab_off + key] = cm.gtab[r.tab_off + key] * ((key & 1u) ? fd : fu);
        } else {
            up[i] = cm.up[i] * fu;
            down[i] = cm.down[i] * fd;
        }
    }
    // Exponent: every non-zero rate an integer multiple of 2^-E where possible,
    // capped so that N * max_rate * 2^E < 2^62 (exact 64-bit sums).
    int e_exact = INT_MIN;
    float vmax = 0.f;
    auto visit = [&](float v) {
        if (v == 0.f) return;
        check_rate(v, "variant");
        int ex;
        const double fr = std::frexp((double)v, &ex);
        const uint32_t mant = (uint32_t)std::ldexp(fr, 24);
        e_exact = std::max(e_exact, -(ex - 24 + __builtin_ctz(mant)));
        vmax = std::max(vmax, v);
    };
    for (int i = 0; i < cm.N; i++)
        if (!cm.generic[i]) {
            visit(up[i]);
            visit(down[i]);
        }
    for (float v : gt) visit(v);
    if (vmax == 0.f) {
        t.E = 0;
    } else {
        int ev;
        std::frexp((double)vmax, &ev);
        t.E = std::min(e_exact, 62 - ev - ceil_log2_u64((uint64_t)cm.N));
    }
    auto fixed = [&](float v) -> uint64_t {
        if (v == 0.f) return 0;
        double r = std::nearbyint(std::ldexp((double)v, t.E));
        return r < 1.0 ? 1ull : (uint64_t)r;
    };
    t.rates.resize(2 * (size_t)cm.N);
    for (int i = 0; i < cm.N; i++) {
        t.rates[2 * i] = fixed(up[i]);
        t.rates[2 * i + 1] = fixed(down[i]);
    }
    t.rtab.resize(gt.size());
    for (size_t k = 0; k < gt.size(); k++) t.rtab[k] = fixed(gt[k]);
    // entropy exponent: the largest possible sum of r log2 r over the readouts stays below 2^62
    double tmax = 0.0;
    for (int j = 0; j < cm.K; j++) {
        const int n = cm.readouts[j];
        double mx = 0.0;
        auto visit_r = [&](uint64_t r) {
            if (!r) return;
            const double rd = (double)r;
            const double v = rd * mbg_log2(rd);
            if (v > mx) mx = v;
        };
        if (cm.generic[n]) {
            const NodeRec& r = cm.nodes[n];
            for (uint32_t key = 0; key < (1u << (r.in_cnt + 1)); key++) visit_r(t.rtab[r.tab_off + key]);
        } else {
            visit_r(t.rates[2 * n]);
            visit_r(t.rates[2 * n + 1]);
        }
        tmax = tmax + mx;
    }
    t.sigma = tmax < 1.0 ? 62 : 62 - ceil_log2_real(tmax);
    t.terms.assign(2 * (size_t)cm.K, 0);
    for (int j = 0; j < cm.K; j++) {
        const int n = cm.readouts[j];
        if (cm.generic[n]) continue;
        for (int b = 0; b < 2; b++) t.terms[2 * j + b] = entropy_term(t.sigma, t.rates[2 * n + b]);
    }
    t.istate.resize(cm.N);
    for (int i = 0; i < cm.N; i++) t.istate[i] = istate ? istate[i] : cm.istate[i];
    return t;
}

}  // namespace mbg
