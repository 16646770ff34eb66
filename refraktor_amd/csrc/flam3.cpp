// flam3 XML import (rfkt::import_flam3, src/librefrakt/src/flame_types.cpp:33-239):
// <flame> attributes give the view (scale is divided by the image height),
// each <xform>/<finalxform> becomes up to three vlinks (pre_ variations on
// an identity affine, the main `coefs` affine with its variations, and the
// `post` affine with a linear variation when no post_ variation exists), the
// palette comes from <color index rgb> entries or a hex <palette> block and is
// stored as HSV.
#include <cctype>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "flame.hpp"

namespace rfk {

namespace {

struct xml_elem {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::string text;
    std::vector<xml_elem> children;
    const std::string* attr(const std::string& k) const {
        for (auto& [n, v] : attrs)
            if (n == k) return &v;
        return nullptr;
    }
};

struct xml_parser {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const std::string& w) { throw std::runtime_error("flam3 xml: " + w); }
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) i++; }
    void skip_misc() {
        while (true) {
            ws();
            if (s.compare(i, 4, "<!--") == 0) {
                size_t e = s.find("-->", i);
                if (e == std::string::npos) fail("unterminated comment");
                i = e + 3;
            } else if (s.compare(i, 2, "<?") == 0 || s.compare(i, 2, "<!") == 0) {
                size_t e = s.find('>', i);
                if (e == std::string::npos) fail("unterminated declaration");
                i = e + 1;
            } else {
                return;
            }
        }
    }
    std::string name() {
        size_t st = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) i++;
        return s.substr(st, i - st);
    }
    xml_elem element() {
        skip_misc();
        if (i >= s.size() || s[i] != '<') fail("expected element");
        i++;
        xml_elem e;
        e.name = name();
        while (true) {
            ws();
            if (i >= s.size()) fail("unterminated tag");
            if (s[i] == '/') {
                if (i + 1 < s.size() && s[i + 1] == '>') { i += 2; return e; }
                fail("bad tag end");
            }
            if (s[i] == '>') { i++; break; }
            std::string an = name();
            ws();
            if (i >= s.size() || s[i] != '=') fail("expected '=' after attribute " + an);
            i++;
            ws();
            char q = s[i];
            if (q != '"' && q != '\'') fail("expected quoted attribute value");
            size_t end = s.find(q, i + 1);
            if (end == std::string::npos) fail("unterminated attribute value");
            e.attrs.emplace_back(an, s.substr(i + 1, end - i - 1));
            i = end + 1;
        }
        // content
        while (true) {
            size_t lt = s.find('<', i);
            if (lt == std::string::npos) fail("unterminated element " + e.name);
            e.text += s.substr(i, lt - i);
            i = lt;
            if (s.compare(i, 4, "<!--") == 0) { skip_misc(); continue; }
            if (s.compare(i, 2, "</") == 0) {
                size_t gt = s.find('>', i);
                if (gt == std::string::npos) fail("unterminated close tag");
                i = gt + 1;
                return e;
            }
            e.children.push_back(element());
        }
    }
};

std::vector<double> doubles(const std::string& s) {
    std::vector<double> out;
    const char* p = s.c_str();
    char* end = nullptr;
    while (*p) {
        double v = std::strtod(p, &end);
        if (end == p) { p++; continue; }
        out.push_back(v);
        p = end;
    }
    return out;
}

double num(const std::string& s) { return std::strtod(s.c_str(), nullptr); }

xform xform_from_xml(const xml_elem& node, const variation_db& db) {
    xform xf;
    vlink vls[3];
    bool has_post = false;
    // longest variation-name prefix of a parameter attribute ("julian_power")
    auto split_param = [&](const std::string& a) -> std::pair<std::string, std::string> {
        std::string best;
        for (auto& [vn, _] : db.variations())
            if (a.size() > vn.size() + 1 && a.compare(0, vn.size(), vn) == 0 && a[vn.size()] == '_' && vn.size() > best.size())
                best = vn;
        if (best.empty()) return {};
        return {best, a.substr(best.size() + 1)};
    };
    for (auto& [raw, val] : node.attrs) {
        std::string an = raw;
        int which = 1;
        if (an.rfind("pre_", 0) == 0) {
            which = 0;
            an = an.substr(4);
            if (an == "blur") an = "gaussian_blur";
        } else if (an.rfind("post_", 0) == 0) {
            which = 2;
        }
        vlink& cur = vls[which];
        if (an == "weight") xf.weight = num(val);
        else if (an == "color") xf.color = num(val);
        else if (an == "color_speed") xf.color_speed = num(val);
        else if (an == "opacity") xf.opacity = num(val);
        else if (an == "coefs" || an == "post") {
            if (an == "post") has_post = true;
            auto v = doubles(val);
            while (v.size() < 6) v.push_back(0.0);
            affine a;
            a.a = v[0]; a.d = v[1]; a.b = v[2]; a.e = v[3]; a.c = v[4]; a.f = v[5];
            vls[an == "coefs" ? 1 : 2].transform = a;
        } else if (db.find(an)) {
            if (!cur.variations.count(an)) cur.variations[an] = make_vardata(db, an);
            cur.variations[an].weight = num(val);
        } else if (auto [vn, pn] = split_param(an); !vn.empty()) {
            if (!cur.variations.count(vn)) cur.variations[vn] = make_vardata(db, vn);
            auto& vd = cur.variations[vn];
            if (vd.params.count(pn)) vd.params[pn] = num(val);
        } else if (an == "animate") {
            if (num(val) == 1.0) {
                vls[1].mod_rotate.call = "increase";
                vls[1].mod_rotate.args["per_loop"] = 360.0;
            }
        }
        // other attributes (chaos, symmetry, ...) are ignored as in the reference
    }
    if (!vls[0].variations.empty()) {
        vls[0].transform = affine::identity();
        xf.vchain.push_back(vls[0]);
    }
    xf.vchain.push_back(vls[1]);
    if (!vls[2].variations.empty() || has_post) {
        if (vls[2].variations.empty()) {
            vls[2].variations["linear"] = make_vardata(db, "linear");
            vls[2].variations["linear"].weight = 1.0;
        }
        xf.vchain.push_back(vls[2]);
    }
    return xf;
}

}  // namespace

flame flame_from_flam3(const std::string& xml, const variation_db& db) {
    xml_parser p{xml};
    xml_elem root = p.element();
    const xml_elem* fn = nullptr;
    if (root.name == "flame") fn = &root;
    else
        for (auto& c : root.children)
            if (c.name == "flame") { fn = &c; break; }
    if (!fn) throw std::runtime_error("flam3 xml: no <flame> element");
    flame f;
    if (auto a = fn->attr("name")) f.name = *a;
    auto size = doubles(fn->attr("size") ? *fn->attr("size") : "0 0");
    if (size.size() < 2 || size[1] == 0) throw std::runtime_error("flam3 xml: bad size attribute");
    f.scale = num(fn->attr("scale") ? *fn->attr("scale") : "1") / size[1];
    f.rotate = fn->attr("rotate") ? num(*fn->attr("rotate")) : 0.0;
    auto center = doubles(fn->attr("center") ? *fn->attr("center") : "0 0");
    center.resize(2, 0.0);
    f.center_x = center[0];
    f.center_y = center[1];
    auto dnum = [&](const char* k) { return fn->attr(k) ? num(*fn->attr(k)) : 0.0; };
    f.gamma = dnum("gamma");
    f.brightness = dnum("brightness");
    f.vibrancy = dnum("vibrancy");
    f.highlight_power = dnum("highlight_power");
    f.gamma_threshold = dnum("gamma_threshold");
    std::vector<std::pair<int, std::string>> chaos_rows;
    int xid = 0;
    for (auto& c : fn->children) {
        if (c.name == "xform") {
            if (auto ch = c.attr("chaos")) chaos_rows.emplace_back(xid, *ch);
            xid++;
            f.xforms.push_back(xform_from_xml(c, db));
        } else if (c.name == "finalxform") {
            f.final_xform = xform_from_xml(c, db);
        }
    }
    if (f.xforms.empty()) throw std::runtime_error("flam3 xml: no xforms");
    const xml_elem* pal = nullptr;
    for (auto& c : fn->children)
        if (c.name == "palette") pal = &c;
    if (pal) {
        auto hexv = [](char ch) -> int {
            if (ch >= '0' && ch <= '9') return ch - '0';
            if (ch >= 'a' && ch <= 'f') return ch - 'a' + 10;
            if (ch >= 'A' && ch <= 'F') return ch - 'A' + 10;
            return -1;
        };
        std::vector<int> digits;
        for (char ch : pal->text)
            if (hexv(ch) >= 0) digits.push_back(hexv(ch));
        for (size_t k = 0; k + 6 <= digits.size(); k += 6) {
            double r = digits[k] * 16 + digits[k + 1], g = digits[k + 2] * 16 + digits[k + 3],
                   b = digits[k + 4] * 16 + digits[k + 5];
            f.palette.push_back(rgb_to_hsv(r / 255.0, g / 255.0, b / 255.0));
        }
    } else {
        size_t count = 0;
        for (auto& c : fn->children)
            if (c.name == "color") count++;
        f.palette.assign(count, {0.0, 0.0, 0.0});
        for (auto& c : fn->children) {
            if (c.name != "color") continue;
            size_t idx = size_t(num(c.attr("index") ? *c.attr("index") : "0"));
            auto rgb = doubles(c.attr("rgb") ? *c.attr("rgb") : "0 0 0");
            rgb.resize(3, 0.0);
            if (idx < f.palette.size()) f.palette[idx] = rgb_to_hsv(rgb[0] / 255.0, rgb[1] / 255.0, rgb[2] / 255.0);
        }
    }
    if (f.palette.size() != 256) throw std::runtime_error("flam3 xml: palette must have 256 colours");
    if (!chaos_rows.empty()) {
        size_t n = f.xforms.size();
        std::vector<std::vector<anima>> table(n, std::vector<anima>(n, anima(1.0)));
        for (auto& [row, txt] : chaos_rows) {
            auto v = doubles(txt);
            for (size_t j = 0; j < n && j < v.size(); j++) table[size_t(row)][j] = v[j];
        }
        f.chaos = table;
    }
    return f;
}

}  // namespace rfk
