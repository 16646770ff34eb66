#include "flang.hpp"

#include <cctype>
#include <functional>
#include <sstream>

namespace rfk::flang {

const char* vtype_name(vtype t) {
    switch (t) {
        case vtype::none: return "none";
        case vtype::integer: return "integer";
        case vtype::decimal: return "decimal";
        case vtype::boolean: return "boolean";
        case vtype::vec2: return "vec2";
        case vtype::vec3: return "vec3";
    }
    return "?";
}

// ---------------------------------------------------------------- lexer

namespace {

enum class tk { ident, integer, decimal, op, end };

struct token {
    tk kind;
    std::string text;
    int line;
    bool space_before;  // whitespace precedes the token
};

std::vector<token> lex(const std::string& s) {
    std::vector<token> out;
    size_t i = 0;
    int line = 1;
    static const char* two_char[] = {"**", "==", "!=", "<=", ">=", "&&", "||", "+=", "-=", "*=", "/=", ":="};
    while (true) {
        bool sp = false;
        while (i < s.size() && std::isspace((unsigned char)s[i])) {
            if (s[i] == '\n') line++;
            i++;
            sp = true;
        }
        if (i >= s.size()) break;
        char c = s[i];
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t st = i;
            while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) i++;
            out.push_back({tk::ident, s.substr(st, i - st), line, sp});
            continue;
        }
        if (std::isdigit((unsigned char)c)) {
            size_t st = i;
            while (i < s.size() && std::isdigit((unsigned char)s[i])) i++;
            bool dec = false;
            if (i + 1 < s.size() && s[i] == '.' && std::isdigit((unsigned char)s[i + 1])) {
                dec = true;
                i++;
                while (i < s.size() && std::isdigit((unsigned char)s[i])) i++;
            }
            out.push_back({dec ? tk::decimal : tk::integer, s.substr(st, i - st), line, sp});
            continue;
        }
        bool matched = false;
        for (const char* op : two_char) {
            if (s.compare(i, 2, op) == 0) {
                out.push_back({tk::op, op, line, sp});
                i += 2;
                matched = true;
                break;
            }
        }
        if (matched) continue;
        if (std::string("+-*/<>=!(){},;.?:").find(c) != std::string::npos) {
            out.push_back({tk::op, std::string(1, c), line, sp});
            i++;
            continue;
        }
        throw syntax_error("line " + std::to_string(line) + ": unexpected character '" + std::string(1, c) + "'");
    }
    out.push_back({tk::end, "", line, true});
    return out;
}

// ---------------------------------------------------------------- parser

struct parser {
    std::vector<token> toks;
    size_t i = 0;

    const token& peek(size_t k = 0) const { return toks[std::min(i + k, toks.size() - 1)]; }
    bool is_op(const char* op, size_t k = 0) const { return peek(k).kind == tk::op && peek(k).text == op; }
    bool is_ident(const char* name, size_t k = 0) const { return peek(k).kind == tk::ident && peek(k).text == name; }

    [[noreturn]] void fail(const std::string& what) const {
        throw syntax_error("line " + std::to_string(peek().line) + ": " + what + " (near '" + peek().text + "')");
    }
    void expect_op(const char* op) {
        if (!is_op(op)) fail(std::string("expected '") + op + "'");
        i++;
    }
    node_ptr make(nk k) {
        auto n = std::make_unique<node>(k);
        n->line = peek().line;
        return n;
    }
    static bool is_keyword(const std::string& s) {
        return s == "if" || s == "else" || s == "break" || s == "true" || s == "false" || s == "and" || s == "or";
    }

    // lvalue: ident | ident '.' ident  (no whitespace around the dot)
    node_ptr lvalue() {
        if (peek().kind != tk::ident || is_keyword(peek().text)) fail("expected a name");
        if (is_op(".", 1) && !peek(1).space_before && peek(2).kind == tk::ident && !peek(2).space_before) {
            auto n = make(nk::member);
            n->text = peek().text;
            n->op = peek(2).text;
            i += 3;
            return n;
        }
        auto n = make(nk::var);
        n->text = peek().text;
        i++;
        return n;
    }

    node_ptr value() {
        const token& t = peek();
        if (t.kind == tk::integer || t.kind == tk::decimal) {
            auto n = make(t.kind == tk::integer ? nk::lit_int : nk::lit_dec);
            n->text = t.text;
            i++;
            return n;
        }
        if (t.kind == tk::op && t.text == "+" && (peek(1).kind == tk::integer || peek(1).kind == tk::decimal) &&
            !peek(1).space_before) {
            // a literal with an explicit sign
            auto n = make(peek(1).kind == tk::integer ? nk::lit_int : nk::lit_dec);
            n->text = peek(1).text;
            i += 2;
            return n;
        }
        if (is_ident("true") || is_ident("false")) {
            auto n = make(nk::lit_bool);
            n->text = t.text;
            i++;
            return n;
        }
        if (is_op("(")) {
            i++;
            auto n = make(nk::paren);
            n->kids.push_back(expression());
            expect_op(")");
            return n;
        }
        if (t.kind == tk::ident) {
            auto lv = lvalue();
            if (is_op("(")) {
                i++;
                auto n = make(nk::call);
                n->line = lv->line;
                n->text = lv->kind == nk::member ? lv->text + "." + lv->op : lv->text;
                if (!is_op(")")) {
                    n->kids.push_back(expression());
                    while (is_op(",")) { i++; n->kids.push_back(expression()); }
                }
                expect_op(")");
                return n;
            }
            return lv;
        }
        fail("expected a value");
    }

    node_ptr unary() {
        if (is_op("-") || is_op("!")) {
            auto n = make(nk::unop);
            n->op = peek().text;
            i++;
            n->kids.push_back(value());
            return n;
        }
        return value();
    }

    using sub_fn = node_ptr (parser::*)();

    node_ptr chain(sub_fn sub, const std::vector<std::string>& ops) {
        node_ptr lhs = (this->*sub)();
        while (true) {
            std::string found;
            for (auto& o : ops) {
                if ((o == "and" || o == "or") ? is_ident(o.c_str()) : is_op(o.c_str())) { found = o; break; }
            }
            if (found.empty()) break;
            i++;
            auto n = make(nk::binop);
            n->op = found == "and" ? "&&" : found == "or" ? "||" : found;
            n->line = lhs->line;
            n->kids.push_back(std::move(lhs));
            n->kids.push_back((this->*sub)());
            lhs = std::move(n);
        }
        return lhs;
    }

    node_ptr power() { return chain(&parser::unary, {"**"}); }
    node_ptr product() { return chain(&parser::power, {"*", "/"}); }
    node_ptr sum() { return chain(&parser::product, {"+", "-"}); }
    node_ptr comparison() { return chain(&parser::sum, {">=", "<=", ">", "<"}); }
    node_ptr equality() { return chain(&parser::comparison, {"==", "!="}); }
    node_ptr boolean() { return chain(&parser::equality, {"&&", "||", "and", "or"}); }

    node_ptr expression() {
        // ternary: '(' boolean ')' '?' boolean ':' boolean
        if (is_op("(")) {
            size_t save = i;
            try {
                i++;
                auto cond = boolean();
                if (is_op(")") && is_op("?", 1)) {
                    i += 2;
                    auto a = boolean();
                    expect_op(":");
                    auto b = boolean();
                    auto n = make(nk::ternary);
                    n->line = cond->line;
                    n->kids.push_back(std::move(cond));
                    n->kids.push_back(std::move(a));
                    n->kids.push_back(std::move(b));
                    return n;
                }
            } catch (const syntax_error&) {
            }
            i = save;
        }
        return boolean();
    }

    node_ptr statement() {
        if (is_ident("if")) {
            auto n = make(nk::if_stmt);
            i++;
            n->kids.push_back(expression());
            n->kids.push_back(statement());
            if (is_ident("else")) {
                i++;
                n->kids.push_back(statement());
            }
            return n;
        }
        if (is_ident("break")) {
            auto n = make(nk::brk);
            i++;
            expect_op(";");
            return n;
        }
        if (is_op("{")) {
            auto n = make(nk::block);
            i++;
            if (is_op("}")) fail("empty block");
            while (!is_op("}")) {
                if (peek().kind == tk::end) fail("unterminated block");
                n->kids.push_back(statement());
            }
            i++;
            return n;
        }
        auto target = lvalue();
        if (is_op(":=")) {
            if (target->kind != nk::var) fail("declarations need a plain name");
            auto n = make(nk::declare);
            n->line = target->line;
            n->text = target->text;
            i++;
            n->kids.push_back(expression());
            expect_op(";");
            return n;
        }
        for (const char* op : {"=", "+=", "-=", "*=", "/="}) {
            if (is_op(op)) {
                auto n = make(nk::assign);
                n->line = target->line;
                n->op = op;
                i++;
                n->kids.push_back(std::move(target));
                n->kids.push_back(expression());
                expect_op(";");
                return n;
            }
        }
        fail("expected an assignment or declaration");
    }
};

}  // namespace

node_ptr parse_statements(const std::string& src) {
    parser p{lex(src)};
    auto root = std::make_unique<node>(nk::block);
    root->root = true;
    if (p.peek().kind == tk::end) throw syntax_error("empty source");
    while (p.peek().kind != tk::end) root->kids.push_back(p.statement());
    return root;
}

node_ptr parse_expression(const std::string& src) {
    parser p{lex(src)};
    auto e = p.expression();
    if (p.peek().kind != tk::end) p.fail("trailing input after expression");
    return e;
}

// ---------------------------------------------------------------- queries

void collect_common(const node* n, std::vector<std::string>& out) {
    if (!n) return;
    if (n->kind == nk::member && n->text == "common") {
        bool seen = false;
        for (auto& s : out) seen |= (s == n->op);
        if (!seen) out.push_back(n->op);
    }
    for (auto& k : n->kids) collect_common(k.get(), out);
}

bool uses_rng(const node* n) {
    if (!n) return false;
    if (n->kind == nk::call && (n->text == "rand01" || n->text == "randbit" || n->text == "randgauss")) return true;
    for (auto& k : n->kids)
        if (uses_rng(k.get())) return true;
    return false;
}

// ---------------------------------------------------------------- checker

namespace {

using sig = std::pair<std::vector<vtype>, vtype>;

const std::map<std::string, std::vector<sig>>& stdlib() {
    static const auto lib = [] {
        std::map<std::string, std::vector<sig>> l;
        using v = vtype;
        for (const char* n : {"sin", "sinpi", "cos", "cospi", "cosh", "sinh", "exp", "sqrt", "eps", "tan", "abs", "log",
                              "log10", "log2", "rsqrt", "acos"})
            l[n] = {{{v::decimal}, v::decimal}, {{v::vec2}, v::vec2}, {{v::vec3}, v::vec3}};
        for (const char* n : {"sincos", "cossin", "sincospi", "cossinpi"}) l[n] = {{{v::decimal}, v::vec2}};
        for (const char* n : {"copysign", "modf", "pow", "atan2", "hypot"}) l[n] = {{{v::decimal, v::decimal}, v::decimal}};
        for (const char* n : {"rand01", "randbit"}) l[n] = {{{}, v::decimal}};
        l["vec2"] = {{{v::decimal, v::decimal}, v::vec2}};
        l["vec3"] = {{{v::decimal, v::decimal, v::decimal}, v::vec3}};
        l["elmul"] = {{{v::vec2, v::vec2}, v::vec2}};
        l["randgauss"] = {{{}, v::vec2}};
        for (const char* n : {"floor", "trunc", "rint"}) l[n] = {{{v::decimal}, v::integer}};
        l["is_even"] = {{{v::integer}, v::boolean}};
        return l;
    }();
    return lib;
}

vtype binop_type(const std::string& op, vtype a, vtype b) {
    using v = vtype;
    auto num = [](v t) { return t == v::integer || t == v::decimal; };
    auto vec = [](v t) { return t == v::vec2 || t == v::vec3; };
    if (op == "+" || op == "-") {
        if (num(a) && num(b)) return (a == v::integer && b == v::integer) ? v::integer : v::decimal;
        if (vec(a) && a == b) return a;
        return v::none;
    }
    if (op == "*") {
        if (num(a) && num(b)) return (a == v::integer && b == v::integer) ? v::integer : v::decimal;
        if (num(a) && vec(b)) return b;
        if (vec(a) && num(b)) return a;
        return v::none;
    }
    if (op == "/") {
        if (num(a) && num(b)) return (a == v::integer && b == v::integer) ? v::integer : v::decimal;
        if (vec(a) && num(b)) return a;
        return v::none;
    }
    if (op == "**") {
        if (num(a) && num(b)) return (a == v::integer && b == v::integer) ? v::integer : v::decimal;
        return v::none;
    }
    if (op == "==" || op == "!=") {
        if (num(a) && num(b)) return v::boolean;
        if (a == v::boolean && b == v::boolean) return v::boolean;
        return v::none;
    }
    if (op == "<" || op == "<=" || op == ">" || op == ">=") return (num(a) && num(b)) ? v::boolean : v::none;
    if (op == "&&" || op == "||") return (a == v::boolean && b == v::boolean) ? v::boolean : v::none;
    return v::none;
}

vtype assign_type(const std::string& op, vtype target, vtype val) {
    using v = vtype;
    auto num = [](v t) { return t == v::integer || t == v::decimal; };
    auto vec = [](v t) { return t == v::vec2 || t == v::vec3; };
    if (op == "=" || op == "+=" || op == "-=") {
        if (num(target) && num(val)) return target;
        if (vec(target) && target == val) return target;
        return v::none;
    }
    // *= and /=
    if (num(target) && num(val)) return target;
    if (vec(target) && num(val)) return target;
    return v::none;
}

struct checker {
    const check_env* env = nullptr;
    const std::set<std::string>* common_only = nullptr;
    std::vector<std::map<std::string, vtype>> scopes;

    [[noreturn]] void fail(const node* n, const std::string& what) {
        throw semantic_error("line " + std::to_string(n->line) + ": " + what);
    }

    const vtype* lookup(const std::string& name) {
        for (auto it = scopes.rbegin(); it != scopes.rend(); ++it) {
            auto f = it->find(name);
            if (f != it->end()) return &f->second;
        }
        return nullptr;
    }

    vtype member_type(const node* n) {
        const std::string& g = n->text;
        const std::string& m = n->op;
        if (g == "math") {
            if (m == "pi" || m == "inv_pi") return vtype::decimal;
            fail(n, "unknown constant math." + m);
        }
        if (g == "common") {
            const auto& names = common_only ? *common_only : env->common;
            if (names.count(m)) return vtype::decimal;
            fail(n, "unknown common value '" + m + "'");
        }
        if (g == "param" && env) {
            if (env->params.count(m)) return vtype::decimal;
            fail(n, "unknown parameter '" + m + "'");
        }
        if (g == "aff" && env) {
            if (m.size() == 1 && std::string("abcdef").find(m[0]) != std::string::npos) return vtype::decimal;
            fail(n, "unknown affine member '" + m + "'");
        }
        const vtype* base = lookup(g);
        if (!base) fail(n, "unknown name '" + g + "'");
        if (*base == vtype::vec2 && (m == "x" || m == "y")) return vtype::decimal;
        if (*base == vtype::vec3 && (m == "x" || m == "y" || m == "z")) return vtype::decimal;
        fail(n, "'" + g + "' has no member '" + m + "'");
    }

    vtype expr(node* n) {
        vtype t = vtype::none;
        switch (n->kind) {
            case nk::lit_int: t = vtype::integer; break;
            case nk::lit_dec: t = vtype::decimal; break;
            case nk::lit_bool: t = vtype::boolean; break;
            case nk::paren: t = expr(n->kid(0)); break;
            case nk::var: {
                const vtype* vt = lookup(n->text);
                if (!vt) fail(n, "unknown name '" + n->text + "'");
                t = *vt;
                break;
            }
            case nk::member: t = member_type(n); break;
            case nk::unop: {
                vtype a = expr(n->kid(0));
                if (n->op == "-") {
                    if (a == vtype::boolean) fail(n, "cannot negate a boolean");
                    t = a;
                } else {
                    if (a != vtype::boolean) fail(n, "'!' needs a boolean");
                    t = a;
                }
                break;
            }
            case nk::binop: {
                vtype a = expr(n->kid(0)), b = expr(n->kid(1));
                t = binop_type(n->op, a, b);
                if (t == vtype::none)
                    fail(n, std::string("no operator ") + n->op + " for " + vtype_name(a) + " and " + vtype_name(b));
                break;
            }
            case nk::ternary: {
                vtype c = expr(n->kid(0)), a = expr(n->kid(1)), b = expr(n->kid(2));
                if (c != vtype::boolean) fail(n, "ternary condition must be boolean");
                if (a == b) t = a;
                else if ((a == vtype::integer || a == vtype::decimal) && (b == vtype::integer || b == vtype::decimal))
                    t = vtype::decimal;
                else fail(n, "ternary branches disagree");
                break;
            }
            case nk::call: {
                auto& lib = stdlib();
                auto f = lib.find(n->text);
                if (f == lib.end()) fail(n, "unknown function '" + n->text + "'");
                std::vector<vtype> args;
                for (auto& k : n->kids) args.push_back(expr(k.get()));
                const sig* match = nullptr;
                for (auto& s : f->second)
                    if (s.first == args) { match = &s; break; }
                if (!match) {
                    for (auto& s : f->second) {
                        if (s.first.size() != args.size()) continue;
                        bool ok = true;
                        for (size_t k = 0; k < args.size(); k++)
                            ok &= (s.first[k] == args[k]) || (s.first[k] == vtype::decimal && args[k] == vtype::integer);
                        if (ok) { match = &s; break; }
                    }
                }
                if (!match) fail(n, "no overload of '" + n->text + "' takes these arguments");
                t = match->second;
                break;
            }
            default: fail(n, "not an expression");
        }
        n->type = t;
        return t;
    }

    vtype target_type(node* n) {
        if (n->kind == nk::var) {
            const vtype* vt = lookup(n->text);
            if (!vt) fail(n, "assignment to undeclared name '" + n->text + "'");
            if (n->text == "p") fail(n, "the input point 'p' is read-only");
            if (n->text == "weight") fail(n, "'weight' is read-only");
            n->type = *vt;
            return *vt;
        }
        if (n->kind == nk::member) {
            if (n->text == "param") {
                if (!env || !env->precalc) fail(n, "parameters can only be written in precalc");
            } else if (n->text == "math" || n->text == "common" || n->text == "aff") {
                fail(n, "'" + n->text + "' is read-only");
            } else if (n->text == "p") {
                fail(n, "the input point 'p' is read-only");
            }
            vtype t = member_type(n);
            n->type = t;
            return t;
        }
        fail(n, "invalid assignment target");
    }

    void stmt(node* n) {
        switch (n->kind) {
            case nk::block:
                scopes.emplace_back();
                for (auto& k : n->kids) stmt(k.get());
                scopes.pop_back();
                break;
            case nk::if_stmt: {
                if (expr(n->kid(0)) != vtype::boolean) fail(n, "if condition must be boolean");
                for (size_t k = 1; k < n->kids.size(); k++) {
                    scopes.emplace_back();
                    stmt(n->kid(k));
                    scopes.pop_back();
                }
                break;
            }
            case nk::assign: {
                vtype tt = target_type(n->kid(0));
                vtype vt = expr(n->kid(1));
                if (assign_type(n->op, tt, vt) == vtype::none)
                    fail(n, std::string("cannot ") + n->op + " a " + vtype_name(vt) + " into a " + vtype_name(tt));
                break;
            }
            case nk::declare: {
                if (scopes.back().count(n->text)) fail(n, "'" + n->text + "' is already declared");
                if (n->text == "p" || n->text == "result" || n->text == "weight")
                    fail(n, "'" + n->text + "' shadows a builtin");
                vtype vt = expr(n->kid(0));
                if (vt == vtype::none || vt == vtype::boolean) fail(n, "cannot declare a value of this type");
                scopes.back()[n->text] = vt;
                break;
            }
            case nk::brk: break;
            default: fail(n, "not a statement");
        }
    }
};

}  // namespace

void check_variation(node* body, const check_env& env) {
    checker c;
    c.env = &env;
    c.scopes.emplace_back();
    c.scopes[0]["p"] = vtype::vec2;
    c.scopes[0]["result"] = vtype::vec2;
    c.scopes[0]["weight"] = vtype::decimal;
    for (auto& k : body->kids) c.stmt(k.get());
}

vtype check_common(node* expr, const std::set<std::string>& common_names) {
    checker c;
    c.common_only = &common_names;
    c.scopes.emplace_back();
    c.scopes[0]["p"] = vtype::vec2;
    return c.expr(expr);
}

// ---------------------------------------------------------------- emitter

namespace {

struct emitter {
    const emit_env& env;

    std::string lval(const node* n) {
        if (n->kind == nk::var) {
            if (n->text == "result") return env.resvar;
            if (n->text == "p") return env.pvar;
            if (n->text == "weight") return env.weight;
            return "u_" + n->text;
        }
        // member
        const std::string& g = n->text;
        const std::string& m = n->op;
        if (g == "param") {
            auto it = env.params.find(m);
            if (it == env.params.end()) throw semantic_error("emit: unknown parameter " + m);
            return it->second;
        }
        if (g == "aff") {
            static const std::string order = "adbecf";
            return "P[" + std::to_string(env.aff_base + int(order.find(m[0]))) + "]";
        }
        if (g == "math") return "fl::" + m;
        if (g == "common") return "common_" + m;
        std::string base = (g == "result") ? env.resvar : (g == "p") ? env.pvar : "u_" + g;
        return base + "." + m;
    }

    static bool is_times(const node* n) { return n->kind == nk::binop && n->op == "*"; }
    static bool scalar_int(vtype t) { return t == vtype::integer; }

    // FMA fusion as flame_compiler.cpp:114-140 does it: (a*b)+c, c+(a*b),
    // (a*b)-c, c-(a*b), x+=a*b, x-=a*b. Pure integer arithmetic is left alone.
    bool fusable(const node* n) const {
        if (n->type == vtype::integer) return false;
        if (n->kind == nk::binop && (n->op == "+" || n->op == "-"))
            return is_times(n->kid(0)) || is_times(n->kid(1));
        return false;
    }

    std::string expr(const node* n) {
        switch (n->kind) {
            case nk::lit_int: return n->text;
            case nk::lit_dec: return "Real(" + n->text + ")";
            case nk::lit_bool: return n->text;
            case nk::paren: return "(" + expr(n->kid(0)) + ")";
            case nk::var:
            case nk::member: return lval(n);
            case nk::unop: return n->op + expr(n->kid(0));
            case nk::ternary:
                return "((" + expr(n->kid(0)) + ") ? (" + expr(n->kid(1)) + ") : (" + expr(n->kid(2)) + "))";
            case nk::binop: {
                if (fusable(n)) {
                    int mc = is_times(n->kid(0)) ? 0 : 1;
                    const node* mul = n->kid(size_t(mc));
                    std::string ml = expr(mul->kid(0)), mr = expr(mul->kid(1)), add = expr(n->kid(size_t(1 - mc)));
                    if (n->op == "+") return "fl::fma(" + ml + ", " + mr + ", " + add + ")";
                    if (mc == 0) return "fl::fma(" + ml + ", " + mr + ", -(" + add + "))";
                    return "fl::fma(" + ml + ", -(" + mr + "), " + add + ")";
                }
                if (n->op == "**") return "fl::pow(Real(" + expr(n->kid(0)) + "), Real(" + expr(n->kid(1)) + "))";
                return expr(n->kid(0)) + " " + n->op + " " + expr(n->kid(1));
            }
            case nk::call: {
                const std::string& f = n->text;
                if (f == "rand01" || f == "randbit" || f == "randgauss") return "rs." + f + "()";
                std::string args;
                for (size_t k = 0; k < n->kids.size(); k++) {
                    if (k) args += ", ";
                    args += expr(n->kid(k));
                }
                return "fl::" + f + "(" + args + ")";
            }
            default: throw semantic_error("emit: not an expression");
        }
    }

    static std::string pad(int indent) { return std::string(size_t(indent) * 4, ' '); }

    static bool has_break(const node* n) {
        if (n->kind == nk::brk) return true;
        for (auto& k : n->kids)
            if (has_break(k.get())) return true;
        return false;
    }

    std::string stmt(const node* n, int ind) {
        switch (n->kind) {
            case nk::block: {
                std::string s = pad(ind) + "{\n";
                for (auto& k : n->kids) s += stmt(k.get(), ind + 1);
                return s + pad(ind) + "}\n";
            }
            case nk::if_stmt: {
                std::string s = pad(ind) + "if (" + expr(n->kid(0)) + ")\n" + branch(n->kid(1), ind);
                if (n->kids.size() == 3) s += pad(ind) + "else\n" + branch(n->kid(2), ind);
                return s;
            }
            case nk::assign: {
                const node* t = n->kid(0);
                const node* v = n->kid(1);
                std::string tgt = lval(t);
                if ((n->op == "+=" || n->op == "-=") && is_times(v) && t->type != vtype::integer) {
                    std::string ml = expr(v->kid(0)), mr = expr(v->kid(1));
                    if (n->op == "+=") return pad(ind) + tgt + " = fl::fma(" + ml + ", " + mr + ", " + tgt + ");\n";
                    return pad(ind) + tgt + " = fl::fma(" + ml + ", -(" + mr + "), " + tgt + ");\n";
                }
                return pad(ind) + tgt + " " + n->op + " " + expr(v) + ";\n";
            }
            case nk::declare: return pad(ind) + "auto u_" + n->text + " = " + expr(n->kid(0)) + ";\n";
            case nk::brk: return pad(ind) + "break;\n";
            default: throw semantic_error("emit: not a statement");
        }
    }

    std::string branch(const node* n, int ind) {
        if (n->kind == nk::block) return stmt(n, ind);
        return stmt(n, ind + 1);
    }
};

}  // namespace

std::string emit_statements(const node* body, const emit_env& env, int indent) {
    emitter e{env};
    bool brk = emitter::has_break(body);
    std::string s;
    std::string inner;
    for (auto& k : body->kids) inner += e.stmt(k.get(), indent + 1);
    if (brk) return std::string(size_t(indent) * 4, ' ') + "do {\n" + inner + std::string(size_t(indent) * 4, ' ') + "} while (0);\n";
    return std::string(size_t(indent) * 4, ' ') + "{\n" + inner + std::string(size_t(indent) * 4, ' ') + "}\n";
}

std::string emit_expression(const node* expr, const emit_env& env) {
    emitter e{env};
    return e.expr(expr);
}

// ---------------------------------------------------------------- printer

std::string to_source(const node* n) {
    switch (n->kind) {
        case nk::block: {
            std::string s = n->root ? "" : "{ ";
            for (auto& k : n->kids) s += to_source(k.get()) + " ";
            if (!n->root) s += "}";
            else if (!s.empty()) s.pop_back();
            return s;
        }
        case nk::if_stmt: {
            std::string s = "if " + to_source(n->kid(0)) + " " + to_source(n->kid(1));
            if (n->kids.size() == 3) s += " else " + to_source(n->kid(2));
            return s;
        }
        case nk::assign: return to_source(n->kid(0)) + " " + n->op + " " + to_source(n->kid(1)) + ";";
        case nk::declare: return n->text + " := " + to_source(n->kid(0)) + ";";
        case nk::brk: return "break;";
        case nk::binop: return "(" + to_source(n->kid(0)) + " " + n->op + " " + to_source(n->kid(1)) + ")";
        case nk::unop: return n->op + to_source(n->kid(0));
        case nk::ternary:
            return "((" + to_source(n->kid(0)) + ") ? " + to_source(n->kid(1)) + " : " + to_source(n->kid(2)) + ")";
        case nk::call: {
            std::string s = n->text + "(";
            for (size_t k = 0; k < n->kids.size(); k++) {
                if (k) s += ", ";
                s += to_source(n->kid(k));
            }
            return s + ")";
        }
        case nk::var: return n->text;
        case nk::member: return n->text + "." + n->op;
        case nk::lit_int:
        case nk::lit_dec:
        case nk::lit_bool: return n->text;
        case nk::paren: return "(" + to_source(n->kid(0)) + ")";
    }
    return "?";
}

}  // namespace rfk::flang
