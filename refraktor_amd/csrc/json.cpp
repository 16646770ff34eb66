#include "json.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace rfk::json {

namespace {

struct parser {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("json parse error at offset " + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { i++; return true; }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }

    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) out.push_back(char(cp));
        else if (cp < 0x800) { out.push_back(char(0xC0 | (cp >> 6))); out.push_back(char(0x80 | (cp & 0x3F))); }
        else if (cp < 0x10000) {
            out.push_back(char(0xE0 | (cp >> 12))); out.push_back(char(0x80 | ((cp >> 6) & 0x3F)));
            out.push_back(char(0x80 | (cp & 0x3F)));
        } else {
            out.push_back(char(0xF0 | (cp >> 18))); out.push_back(char(0x80 | ((cp >> 12) & 0x3F)));
            out.push_back(char(0x80 | ((cp >> 6) & 0x3F))); out.push_back(char(0x80 | (cp & 0x3F)));
        }
    }

    std::string string_body() {
        // opening quote already consumed
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            char c = s[i++];
            if (c == '"') break;
            if (c != '\\') { out.push_back(c); continue; }
            if (i >= s.size()) fail("bad escape");
            char e = s[i++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    if (i + 4 > s.size()) fail("bad \\u escape");
                    uint32_t cp = uint32_t(std::strtoul(s.substr(i, 4).c_str(), nullptr, 16));
                    i += 4;
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }

    value parse_value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            i++;
            value v = value::make_object();
            if (eat('}')) return v;
            do {
                ws();
                if (i >= s.size() || s[i] != '"') fail("expected key");
                i++;
                std::string key = string_body();
                expect(':');
                (*v.obj)[key] = parse_value();
            } while (eat(','));
            expect('}');
            return v;
        }
        if (c == '[') {
            i++;
            value v = value::make_array();
            if (eat(']')) return v;
            do { v.arr->push_back(parse_value()); } while (eat(','));
            expect(']');
            return v;
        }
        if (c == '"') { i++; return value::make_string(string_body()); }
        if (s.compare(i, 4, "true") == 0) { i += 4; return value::make_bool(true); }
        if (s.compare(i, 5, "false") == 0) { i += 5; return value::make_bool(false); }
        if (s.compare(i, 4, "null") == 0) { i += 4; return value{}; }
        // number
        size_t start = i;
        if (s[i] == '-' || s[i] == '+') i++;
        bool integral = true;
        while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' || s[i] == 'e' || s[i] == 'E' ||
                                ((s[i] == '-' || s[i] == '+') && (s[i - 1] == 'e' || s[i - 1] == 'E')))) {
            if (s[i] == '.' || s[i] == 'e' || s[i] == 'E') integral = false;
            i++;
        }
        if (start == i) fail("unexpected character");
        std::string tok = s.substr(start, i - start);
        char* end = nullptr;
        double d = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str()) fail("bad number");
        return value::make_number(d, integral);
    }
};

void dump_to(std::ostringstream& os, const value& v) {
    switch (v.k) {
        case value::kind::null: os << "null"; break;
        case value::kind::boolean: os << (v.b ? "true" : "false"); break;
        case value::kind::number: {
            char buf[64];
            if (v.is_integer && std::fabs(v.num) < 9.0e15) std::snprintf(buf, sizeof buf, "%lld", (long long)v.num);
            else std::snprintf(buf, sizeof buf, "%.17g", v.num);
            os << buf;
            break;
        }
        case value::kind::string: {
            os << '"';
            for (char c : v.str) {
                if (c == '"' || c == '\\') os << '\\' << c;
                else if (c == '\n') os << "\\n";
                else os << c;
            }
            os << '"';
            break;
        }
        case value::kind::array: {
            os << '[';
            bool first = true;
            for (auto& e : *v.arr) { if (!first) os << ','; first = false; dump_to(os, e); }
            os << ']';
            break;
        }
        case value::kind::object: {
            os << '{';
            bool first = true;
            for (auto& [k, e] : *v.obj) {
                if (!first) os << ',';
                first = false;
                dump_to(os, value::make_string(k));
                os << ':';
                dump_to(os, e);
            }
            os << '}';
            break;
        }
    }
}

}  // namespace

value parse(const std::string& text) {
    parser p{text};
    value v = p.parse_value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

std::string dump(const value& v) {
    std::ostringstream os;
    dump_to(os, v);
    return os.str();
}

}  // namespace rfk::json
