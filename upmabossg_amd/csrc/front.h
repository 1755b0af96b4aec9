// Front end of the MaBoSS .bnd/.cfg language, hand-written (the reference
// generates its scanner and parser with Flex/Bison: src/parser/scanner.ll,
// src/parser/parser.yy; neither tool is required here).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mbg {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Expression tree (src/parser/expressions.h:11-120).
struct Ast {
    enum Kind { Number, Node, Variable, Alias, Unary, Binary, Ternary, Group };
    enum Oper { Add, Sub, Mul, Div, And, Or, Xor, Not, Le, Lt, Ge, Gt, Eq, Ne, Pos };
    Kind kind;
    Oper oper = Add;
    float number = 0.f;
    std::string text;  // node / $variable / @alias name
    std::vector<std::unique_ptr<Ast>> kids;
    int height = 1;  // 1 + the tallest kid (bounded by the parser: later passes recurse)
};
using AstPtr = std::unique_ptr<Ast>;

struct NodeDecl {
    std::string name;
    std::vector<std::pair<std::string, AstPtr>> attrs;  // in declaration order
    float istate = 0.5f;                                // src/parser/parse_types.cpp:3

    // First attribute of that name wins (src/parser/parse_types.cpp:6-13).
    const Ast* attr(const std::string& n) const {
        for (const auto& a : attrs)
            if (a.first == n) return a.second.get();
        return nullptr;
    }
};

// Parsed model: the state of the reference's `driver` after parsing
// (src/parser/driver.h:38-43, defaults src/parser/driver.cpp:7-14).
struct ParsedModel {
    std::map<std::string, float> variables;
    std::map<std::string, float> constants{{"discrete_time", 0.f},
                                           {"sample_count", 1000000.f},
                                           {"max_time", 10.f},
                                           {"time_tick", 0.1f},
                                           {"seed_pseudorandom", 0.f}};
    std::vector<NodeDecl> nodes;

    int index_of(const std::string& n) const;
    float constant(const std::string& n) const { return constants.at(n); }
};

// Parse the .bnd text, then the .cfg text, into `out` (driver::parse, src/parser/driver.cpp:16-31).
void parse_model(const std::string& bnd, const std::string& cfg, ParsedModel& out);

// Constant evaluation (expression::evaluate, src/parser/expressions.cpp).
float evaluate(const Ast* e, const ParsedModel& m);

}  // namespace mbg
