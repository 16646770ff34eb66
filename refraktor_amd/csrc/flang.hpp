// The variation language ("flang") used by the variation database: parser,
// type checker and HIP source emitter.
//
// Grammar follows src/librefrakt/depend/flang/src/grammar_defs.h:1-112
// (statements: if/else, `lhs op= expr;`, `name := expr;`, `{...}`, `break;`;
// expression precedence ternary < boolean < equality < comparison < sum <
// product < power < unary < value), with binary operator chains folded
// left-associatively as the reference's `rearrange` transform does
// (depend/flang/src/ast.cpp:8-37). Operator typing follows
// depend/flang/src/grammar.cpp:3-140 and the standard library signatures
// follow grammar.cpp:498-587.
#pragma once

#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace rfk::flang {

enum class vtype { none, integer, decimal, boolean, vec2, vec3 };
const char* vtype_name(vtype t);

enum class nk {
    block,      // statements; `root` flag marks the top level
    if_stmt,    // kids: cond, then, [else]
    assign,     // op: = += -= *= /=; kids: target, value
    declare,    // text: name; kids: value
    brk,        // break;
    binop,      // op; kids: lhs, rhs
    unop,       // op: - or !; kids: operand
    ternary,    // kids: cond, a, b
    call,       // text: function name; kids: args
    var,        // text: name
    member,     // text: base name, op: member name
    lit_int,    // text
    lit_dec,    // text
    lit_bool,   // text
    paren,      // kids: inner
};

struct node {
    nk kind;
    std::string text;
    std::string op;
    std::vector<std::unique_ptr<node>> kids;
    int line = 0;
    bool root = false;
    vtype type = vtype::none;  // filled by the checker for expressions

    node(nk k) : kind(k) {}
    const node* kid(size_t i) const { return kids.at(i).get(); }
    node* kid(size_t i) { return kids.at(i).get(); }
};

using node_ptr = std::unique_ptr<node>;

struct syntax_error : std::runtime_error { using std::runtime_error::runtime_error; };
struct semantic_error : std::runtime_error { using std::runtime_error::runtime_error; };

// A whole statement list (a variation body or a precalc body).
node_ptr parse_statements(const std::string& src);
// A single expression (a `common` definition).
node_ptr parse_expression(const std::string& src);

// Names of `common.X` members referenced anywhere under n.
void collect_common(const node* n, std::vector<std::string>& out);
bool uses_rng(const node* n);

// Type checking environment for one variation.
struct check_env {
    std::set<std::string> common;      // available common names
    std::set<std::string> params;      // param group members
    bool precalc = false;              // precalc bodies may assign params
};
// Throws semantic_error.
void check_variation(node* body, const check_env& env);
vtype check_common(node* expr, const std::set<std::string>& common_names);

// Emission context: how the variation's names map onto generated HIP.
struct emit_env {
    std::string weight;                              // e.g. "P[17]"
    std::map<std::string, std::string> params;       // param name -> lvalue text
    std::string aff;                                 // affine base, e.g. "P[11]"; members are aff + idx
    int aff_base = -1;                               // offset of the vlink affine in P
    std::string pvar = "p";                          // input point
    std::string resvar = "res";                      // accumulated output
};

// Emit a statement list as HIP C++ (wrapped in a do/while when it breaks).
std::string emit_statements(const node* body, const emit_env& env, int indent);
std::string emit_expression(const node* expr, const emit_env& env);

// Debug printing of the AST in a normalised form (round-trip tests).
std::string to_source(const node* n);

}  // namespace rfk::flang
