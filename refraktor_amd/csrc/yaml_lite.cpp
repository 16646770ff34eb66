#include "yaml_lite.hpp"

#include <fstream>
#include <sstream>

namespace rfk::yaml {

namespace {

struct line_t {
    int indent;        // leading spaces
    std::string text;  // content without leading spaces, comment not stripped
    int number;        // 1-based line number
    bool blank;
};

std::string rtrim(std::string s) {
    while (!s.empty() && (s.back() == ' ' || s.back() == '\t' || s.back() == '\r')) s.pop_back();
    return s;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) b--;
    return s.substr(a, b - a);
}

// strip a trailing comment that is outside quotes and preceded by whitespace
std::string strip_comment(const std::string& s) {
    char quote = 0;
    for (size_t i = 0; i < s.size(); i++) {
        char c = s[i];
        if (quote) {
            if (c == quote) quote = 0;
            continue;
        }
        if (c == '"' || c == '\'') quote = c;
        else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return rtrim(s.substr(0, i));
    }
    return s;
}

struct flow_parser {
    const std::string& s;
    size_t i = 0;
    int line;

    [[noreturn]] void fail(const std::string& what) {
        throw parse_error("yaml line " + std::to_string(line) + ": " + what);
    }
    void ws() { while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++; }

    std::string scalar_token() {
        ws();
        if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
            char q = s[i++];
            std::string out;
            while (i < s.size() && s[i] != q) {
                if (q == '"' && s[i] == '\\' && i + 1 < s.size()) { out.push_back(s[i + 1]); i += 2; continue; }
                out.push_back(s[i++]);
            }
            if (i >= s.size()) fail("unterminated quote");
            i++;
            return out;
        }
        size_t st = i;
        while (i < s.size() && s[i] != ',' && s[i] != ']' && s[i] != '}' && !(s[i] == ':' && (i + 1 >= s.size() || s[i + 1] == ' ')))
            i++;
        return trim(s.substr(st, i - st));
    }

    node value() {
        ws();
        if (i >= s.size()) fail("unexpected end of flow value");
        if (s[i] == '[') {
            i++;
            node n; n.k = node::kind::sequence;
            ws();
            if (i < s.size() && s[i] == ']') { i++; return n; }
            while (true) {
                n.items.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; break; }
                fail("expected ',' or ']'");
            }
            return n;
        }
        if (s[i] == '{') {
            i++;
            node n; n.k = node::kind::mapping;
            ws();
            if (i < s.size() && s[i] == '}') { i++; return n; }
            while (true) {
                std::string key = scalar_token();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':' in flow mapping");
                i++;
                n.pairs.emplace_back(key, value());
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; break; }
                fail("expected ',' or '}'");
            }
            return n;
        }
        node n; n.k = node::kind::scalar;
        n.scalar = scalar_token();
        return n;
    }
};

struct block_parser {
    std::vector<line_t> lines;
    size_t pos = 0;

    [[noreturn]] void fail(int line, const std::string& what) {
        throw parse_error("yaml line " + std::to_string(line) + ": " + what);
    }

    void skip_blank() {
        while (pos < lines.size() && lines[pos].blank) pos++;
    }

    // parse a literal block scalar whose header sits on line `header_line`
    node block_scalar(int parent_indent, bool strip, int header_line) {
        std::vector<std::string> out;
        int block_indent = -1;
        while (pos < lines.size()) {
            const auto& ln = lines[pos];
            if (ln.blank) { out.emplace_back(""); pos++; continue; }
            if (ln.indent <= parent_indent) break;
            if (block_indent < 0) block_indent = ln.indent;
            if (ln.indent < block_indent) fail(ln.number, "bad indentation in block scalar");
            out.push_back(std::string(size_t(ln.indent - block_indent), ' ') + rtrim(ln.text));
            pos++;
        }
        (void)header_line;
        // trailing blank lines belong to the chomping indicator
        while (!out.empty() && out.back().empty()) out.pop_back();
        std::string joined;
        for (size_t k = 0; k < out.size(); k++) {
            joined += out[k];
            if (k + 1 < out.size()) joined += "\n";
        }
        if (!strip && !joined.empty()) joined += "\n";
        node n; n.k = node::kind::scalar; n.scalar = joined;
        return n;
    }

    node inline_value(const std::string& raw, int indent, int number) {
        std::string v = trim(strip_comment(raw));
        if (v == "|" || v == "|-" || v == "|+") {
            return block_scalar(indent, v == "|-", number);
        }
        if (v.empty()) {
            // nested block (mapping or sequence) or null
            skip_blank();
            if (pos < lines.size() && lines[pos].indent > indent) return block(lines[pos].indent);
            if (pos < lines.size() && lines[pos].indent == indent && lines[pos].text.rfind("- ", 0) == 0)
                return block(indent);
            return node{};
        }
        if (v[0] == '[' || v[0] == '{') {
            flow_parser fp{v, 0, number};
            node n = fp.value();
            return n;
        }
        node n; n.k = node::kind::scalar;
        if ((v.front() == '"' || v.front() == '\'') && v.size() >= 2 && v.back() == v.front()) {
            flow_parser fp{v, 0, number};
            n.scalar = fp.scalar_token();
        } else {
            // plain scalar, possibly continued on more-indented lines (folded with spaces)
            std::string acc = v;
            while (pos < lines.size() && !lines[pos].blank && lines[pos].indent > indent) {
                acc += " " + trim(strip_comment(lines[pos].text));
                pos++;
            }
            n.scalar = acc;
        }
        return n;
    }

    static size_t key_colon(const std::string& t) {
        char quote = 0;
        for (size_t i = 0; i < t.size(); i++) {
            char c = t[i];
            if (quote) { if (c == quote) quote = 0; continue; }
            if (c == '"' || c == '\'') { quote = c; continue; }
            if (c == '[' || c == '{') return std::string::npos;
            if (c == ':' && (i + 1 == t.size() || t[i + 1] == ' ' || t[i + 1] == '\t')) return i;
        }
        return std::string::npos;
    }

    node block(int indent) {
        skip_blank();
        node n;
        if (pos >= lines.size()) return n;
        if (lines[pos].text.rfind("- ", 0) == 0 || lines[pos].text == "-") {
            n.k = node::kind::sequence;
            while (pos < lines.size()) {
                skip_blank();
                if (pos >= lines.size()) break;
                const auto ln = lines[pos];
                if (ln.indent < indent) break;
                if (ln.indent > indent) fail(ln.number, "bad indentation in sequence");
                if (!(ln.text.rfind("- ", 0) == 0 || ln.text == "-")) break;
                pos++;
                std::string rest = ln.text.size() > 2 ? ln.text.substr(2) : std::string();
                n.items.push_back(inline_value(rest, indent, ln.number));
            }
            return n;
        }
        n.k = node::kind::mapping;
        while (pos < lines.size()) {
            skip_blank();
            if (pos >= lines.size()) break;
            const auto ln = lines[pos];
            if (ln.indent < indent) break;
            if (ln.indent > indent) fail(ln.number, "unexpected indentation");
            size_t c = key_colon(ln.text);
            if (c == std::string::npos) fail(ln.number, "expected 'key: value'");
            std::string key = trim(ln.text.substr(0, c));
            if (key.size() >= 2 && (key.front() == '"' || key.front() == '\'')) key = key.substr(1, key.size() - 2);
            pos++;
            n.pairs.emplace_back(key, inline_value(ln.text.substr(c + 1), indent, ln.number));
        }
        return n;
    }
};

}  // namespace

node parse(const std::string& text) {
    block_parser bp;
    std::istringstream is(text);
    std::string raw;
    int number = 0;
    while (std::getline(is, raw)) {
        number++;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        int ind = 0;
        while (ind < int(raw.size()) && raw[size_t(ind)] == ' ') ind++;
        std::string rest = raw.substr(size_t(ind));
        bool blank = trim(rest).empty() || trim(rest)[0] == '#';
        bp.lines.push_back({ind, rest, number, blank});
    }
    node root = bp.block(0);
    bp.skip_blank();
    if (bp.pos < bp.lines.size()) bp.fail(bp.lines[bp.pos].number, "trailing content");
    return root;
}

node parse_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw parse_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace rfk::yaml
