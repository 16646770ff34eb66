#include "flame.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace rfk {

namespace {
constexpr double kPi = 3.14159265358979323846;

double arg_double(const std::map<std::string, anima::arg_t>& args, const std::string& key, double dflt) {
    auto it = args.find(key);
    if (it == args.end()) return dflt;
    if (auto d = std::get_if<double>(&it->second)) return *d;
    if (auto i = std::get_if<int>(&it->second)) return double(*i);
    return dflt;
}
bool arg_bool(const std::map<std::string, anima::arg_t>& args, const std::string& key, bool dflt) {
    auto it = args.find(key);
    if (it == args.end()) return dflt;
    if (auto b = std::get_if<bool>(&it->second)) return *b;
    return dflt;
}

// Built-in animation functions. `increase` and `sine` are the two functions
// the reference front ends register (src/refraktor/src/refraktor.cpp:351-367),
// `mix` the built-in smoothstep cross-fade (src/librefrakt/src/anima.cpp:35-105).
double call_function(const std::string& name, double t, double iv, const std::map<std::string, anima::arg_t>& args,
                     const std::string& prefix) {
    if (name == "increase") return iv + t * arg_double(args, prefix + "per_loop", 360.0);
    if (name == "sine") {
        double v = std::sin(t * arg_double(args, prefix + "frequency", 1.0) * kPi * 2.0 +
                            arg_double(args, prefix + "phase", 0.0) * kPi / 180.0);
        double sharp = arg_double(args, prefix + "sharpness", 0.0);
        if (sharp > 0) v = std::copysign(1.0, v) * std::pow(std::fabs(v), sharp);
        if (arg_bool(args, prefix + "absolute", false)) v = std::fabs(v);
        return iv + v * arg_double(args, prefix + "amplitude", 1.0);
    }
    if (name == "mix") {
        double left = iv;
        if (auto it = args.find(prefix + "left.function"); it != args.end()) {
            if (auto s = std::get_if<std::string>(&it->second)) left = call_function(*s, t, iv, args, prefix + "left.");
        }
        // a missing or non-numeric right value, start time or length makes
        // the reference log an error and return the initial value
        auto number = [&](const std::string& key, double& out) {
            auto it = args.find(prefix + key);
            if (it == args.end()) return false;
            if (auto d = std::get_if<double>(&it->second)) out = *d;
            else if (auto i = std::get_if<int>(&it->second)) out = double(*i);
            else return false;
            return true;
        };
        double right = 0.0, start = 0.0, length = 0.0;
        if (!number("right.value", right)) return iv;
        if (auto it = args.find(prefix + "right.function"); it != args.end()) {
            if (auto s = std::get_if<std::string>(&it->second))
                right = call_function(*s, t, right, args, prefix + "right.");
        }
        if (!number("start_time", start) || !number("length", length)) return iv;
        if (t <= start) return left;
        if (t >= start + length) return right;
        double m = (t - start) / length;
        m = 6 * m * m * m * m * m - 15 * m * m * m * m + 10 * m * m * m;
        return left * (1 - m) + right * m;
    }
    return iv;  // unknown function: the initial value, as the reference logs and returns iv
}
}  // namespace

double anima::sample(double t) const {
    if (!call) return t0;
    return call_function(*call, t, t0, args, "");
}

// glm::rotate(m, -radians(deg), z) on the 2x2 part (flame_types.h:120-135)
affine_v affine_v::rotated(double deg) const {
    double rad = -deg * kPi / 180.0;
    double cs = std::cos(rad), sn = std::sin(rad);
    return {a * cs + d * sn, -a * sn + d * cs, b * cs + e * sn, -b * sn + e * cs, c, f};
}

size_t flame::size_reals() const {
    size_t s = final_xform ? final_xform->size_reals() : 0;
    if (chaos) s += xforms.size() * (xforms.size() + 1);
    for (auto& x : xforms) s += x.size_reals();
    return s + 13;
}

std::vector<size_t> flame::affine_indices() const {
    std::vector<size_t> r{0, 6};
    size_t idx = chaos ? xforms.size() * (xforms.size() + 1) : 0;
    idx += 13;
    auto walk = [&](const xform& xf) {
        idx += 4;
        for (auto& vl : xf.vchain) {
            r.push_back(idx);
            idx += vl.size_reals();
        }
    };
    for (auto& xf : xforms) walk(xf);
    if (final_xform) walk(*final_xform);
    return r;
}

affine_v flame::screen_space(int w, int h, double t) const {
    affine_v id{1, 0, 0, 1, 0, 0};
    return id.translated(w / 2.0, h / 2.0)
        .scaled(scale.sample(t) * h)
        .rotated(rotate.sample(t))
        .translated(-center_x.sample(t), -center_y.sample(t));
}

affine_v flame::plane_space(int w, int h, double t) const {
    affine_v id{1, 0, 0, 1, 0, 0};
    return id.translated(center_x.sample(t), center_y.sample(t))
        .rotated(-rotate.sample(t))
        .scaled(1 / (scale.sample(t) * h))
        .translated(w / -2.0, h / -2.0);
}

void flame::pack_sample(std::vector<double>& p, double t, int w, int h) const {
    auto put_aff = [&](const affine_v& a) {
        p.push_back(a.a); p.push_back(a.d); p.push_back(a.b);
        p.push_back(a.e); p.push_back(a.c); p.push_back(a.f);
    };
    put_aff(screen_space(w, h, t));
    put_aff(plane_space(w, h, t));
    p.push_back(0.0);  // weight sum, filled by precalc
    if (chaos) {
        for (auto& row : *chaos) {
            p.push_back(0.0);
            for (auto& col : row) p.push_back(col.sample(t));
        }
    }
    auto put_xform = [&](const xform& xf) {
        p.push_back(xf.weight.sample(t));
        p.push_back(xf.color.sample(t));
        p.push_back(xf.color_speed.sample(t));
        p.push_back(xf.opacity.sample(t));
        for (auto& vl : xf.vchain) {
            affine_v a{vl.transform.a.sample(t), vl.transform.d.sample(t), vl.transform.b.sample(t),
                       vl.transform.e.sample(t), vl.transform.c.sample(t), vl.transform.f.sample(t)};
            put_aff(a.scaled(vl.mod_scale.sample(t)).rotated(vl.mod_rotate.sample(t))
                        .translated(vl.mod_x.sample(t), vl.mod_y.sample(t)));
            for (auto& [_, vd] : vl.variations) {
                p.push_back(vd.weight.sample(t));
                for (auto& [__, pv] : vd.params) p.push_back(pv.sample(t));
                for (size_t k = 0; k < vd.precalc_count; k++) p.push_back(0.0);
            }
        }
    };
    for (auto& xf : xforms) put_xform(xf);
    if (final_xform) put_xform(*final_xform);
    for (auto& [hue, sat, val] : palette) {
        p.push_back(std::fmod(hue + mod_hue.sample(t), 360.0));
        p.push_back(std::clamp(sat + mod_sat.sample(t), 0.0, 1.0));
        p.push_back(std::clamp(val + mod_val.sample(t), 0.0, 1.0));
    }
}

flame_layout make_layout(const flame& f, const variation_db& db) {
    flame_layout L;
    L.size_reals = int(f.size_reals());
    L.affine_indices = f.affine_indices();
    int idx = 13;
    if (f.chaos) {
        L.chaos_base = idx;
        idx += int(f.xforms.size() * (f.xforms.size() + 1));
    }
    auto walk = [&](const xform& xf) {
        xform_slot xs;
        xs.weight = idx++;
        xs.color = idx++;
        xs.color_speed = idx++;
        xs.opacity = idx++;
        for (auto& vl : xf.vchain) {
            vlink_slot vs;
            vs.aff = idx;
            idx += 6;
            for (auto& [name, vd] : vl.variations) {
                var_slot v;
                v.name = name;
                v.weight = idx++;
                for (auto& [pn, _] : vd.params) v.user.emplace_back(pn, idx++);
                const variation_def* def = db.find(name);
                if (!def) throw std::runtime_error("unknown variation " + name);
                auto pre = def->precalc_params();
                if (pre.size() != vd.precalc_count) throw std::runtime_error("precalc slot mismatch for " + name);
                for (auto& pn : pre) v.precalc.emplace_back(pn, idx++);
                vs.vars.push_back(std::move(v));
            }
            xs.vlinks.push_back(std::move(vs));
        }
        return xs;
    };
    for (auto& xf : f.xforms) L.xforms.push_back(walk(xf));
    if (f.final_xform) L.final_xform = walk(*f.final_xform);
    if (idx != L.size_reals) throw std::runtime_error("layout size mismatch");
    return L;
}

vardata make_vardata(const variation_db& db, const std::string& name) {
    const variation_def* def = db.find(name);
    if (!def) throw std::runtime_error("unknown variation `" + name + "`");
    vardata v;
    v.weight = 0.0;
    for (auto& [pn, pd] : def->params) {
        if (pd.is_precalc()) {
            v.precalc_count++;
            continue;
        }
        v.params[pn] = pd.default_value.value_or(0.0);
    }
    return v;
}

// ------------------------------------------------------------------ JSON

namespace {

anima anima_from(const json::value& js) {
    if (js.is_number()) return anima(js.num);
    if (!js.is_object() || !js.contains("t0") || !js.at("t0").is_number()) throw std::runtime_error("bad anima");
    anima a(js.at("t0").num);
    if (js.contains("call")) {
        a.call = js.at("call").str;
        if (js.contains("args") && js.at("args").is_object()) {
            for (auto& [k, v] : *js.at("args").obj) {
                if (v.is_number()) {
                    if (v.is_integer) a.args[k] = int(v.num);
                    else a.args[k] = v.num;
                } else if (v.is_bool()) a.args[k] = v.b;
                else if (v.is_string()) a.args[k] = v.str;
            }
        }
    }
    return a;
}

json::value anima_to(const anima& a) {
    if (!a.call) return json::value::make_number(a.t0);
    auto o = json::value::make_object();
    (*o.obj)["t0"] = json::value::make_number(a.t0);
    (*o.obj)["call"] = json::value::make_string(*a.call);
    auto args = json::value::make_object();
    for (auto& [k, v] : a.args) {
        if (auto i = std::get_if<int>(&v)) (*args.obj)[k] = json::value::make_number(*i, true);
        else if (auto d = std::get_if<double>(&v)) (*args.obj)[k] = json::value::make_number(*d);
        else if (auto b = std::get_if<bool>(&v)) (*args.obj)[k] = json::value::make_bool(*b);
        else (*args.obj)[k] = json::value::make_string(std::get<std::string>(v));
    }
    (*o.obj)["args"] = args;
    return o;
}

#define ANIMA_FIELD(obj, js, field)                                                      \
    do {                                                                                 \
        if (!(js).contains(#field)) throw std::runtime_error("missing field " #field);   \
        (obj).field = anima_from((js).at(#field));                                       \
    } while (0)

vardata vardata_from(const std::string& name, const json::value& js, const variation_db& db) {
    vardata vd = make_vardata(db, name);
    if (js.is_number() || (js.is_object() && js.contains("t0"))) {
        if (!vd.params.empty()) throw std::runtime_error("variation " + name + " needs an object with parameters");
        vd.weight = anima_from(js);
        return vd;
    }
    if (!js.is_object() || !js.contains("weight")) throw std::runtime_error("bad variation data for " + name);
    vd.weight = anima_from(js.at("weight"));
    if (js.contains("parameters")) {
        const auto& ps = js.at("parameters");
        if (!ps.is_object()) throw std::runtime_error("parameters must be an object");
        for (auto& [k, v] : *ps.obj) {
            if (!vd.params.count(k)) throw std::runtime_error("unknown parameter " + name + "." + k);
            vd.params[k] = anima_from(v);
        }
    }
    return vd;
}

affine affine_from(const json::value& js) {
    if (!js.is_array() || js.arr->size() != 6) throw std::runtime_error("affine must have 6 entries");
    const auto& a = *js.arr;
    affine r;
    r.a = anima_from(a[0]); r.d = anima_from(a[1]); r.b = anima_from(a[2]);
    r.e = anima_from(a[3]); r.c = anima_from(a[4]); r.f = anima_from(a[5]);
    return r;
}

vlink vlink_from(const json::value& js, const variation_db& db) {
    if (!js.is_object()) throw std::runtime_error("vlink must be an object");
    vlink vl;
    vl.transform = affine_from(js.at("transform"));
    ANIMA_FIELD(vl, js, mod_x);
    ANIMA_FIELD(vl, js, mod_y);
    ANIMA_FIELD(vl, js, mod_scale);
    ANIMA_FIELD(vl, js, mod_rotate);
    if (!js.contains("variations") || !js.at("variations").is_object()) throw std::runtime_error("vlink needs variations");
    for (auto& [k, v] : *js.at("variations").obj) vl.variations[k] = vardata_from(k, v, db);
    return vl;
}

xform xform_from(const json::value& js, const variation_db& db) {
    if (!js.is_object()) throw std::runtime_error("xform must be an object");
    xform xf;
    ANIMA_FIELD(xf, js, weight);
    ANIMA_FIELD(xf, js, color);
    ANIMA_FIELD(xf, js, color_speed);
    ANIMA_FIELD(xf, js, opacity);
    if (!js.contains("vchain") || !js.at("vchain").is_array()) throw std::runtime_error("xform needs a vchain");
    for (auto& v : *js.at("vchain").arr) xf.vchain.push_back(vlink_from(v, db));
    if (xf.vchain.empty()) throw std::runtime_error("xform vchain is empty");
    return xf;
}

json::value xform_to(const xform& xf) {
    auto o = json::value::make_object();
    (*o.obj)["weight"] = anima_to(xf.weight);
    (*o.obj)["color"] = anima_to(xf.color);
    (*o.obj)["color_speed"] = anima_to(xf.color_speed);
    (*o.obj)["opacity"] = anima_to(xf.opacity);
    auto vc = json::value::make_array();
    for (auto& vl : xf.vchain) {
        auto v = json::value::make_object();
        auto t = json::value::make_array();
        for (auto* a : {&vl.transform.a, &vl.transform.d, &vl.transform.b, &vl.transform.e, &vl.transform.c,
                        &vl.transform.f})
            t.arr->push_back(anima_to(*a));
        (*v.obj)["transform"] = t;
        (*v.obj)["mod_x"] = anima_to(vl.mod_x);
        (*v.obj)["mod_y"] = anima_to(vl.mod_y);
        (*v.obj)["mod_scale"] = anima_to(vl.mod_scale);
        (*v.obj)["mod_rotate"] = anima_to(vl.mod_rotate);
        auto vars = json::value::make_object();
        for (auto& [name, vd] : vl.variations) {
            if (vd.params.empty()) {
                (*vars.obj)[name] = anima_to(vd.weight);
            } else {
                auto e = json::value::make_object();
                (*e.obj)["weight"] = anima_to(vd.weight);
                auto ps = json::value::make_object();
                for (auto& [pn, pv] : vd.params) (*ps.obj)[pn] = anima_to(pv);
                (*e.obj)["parameters"] = ps;
                (*vars.obj)[name] = e;
            }
        }
        (*v.obj)["variations"] = vars;
        vc.arr->push_back(v);
    }
    (*o.obj)["vchain"] = vc;
    return o;
}

// base64url without padding (cppcodec base64_url_unpadded) + the
// size-prefixed zlib stream of rfkt::zlib::compress (util/zlib.cpp:9-47).
std::vector<unsigned char> b64url_decode(const std::string& s) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '-' || c == '+') return 62;
        if (c == '_' || c == '/') return 63;
        return -1;
    };
    std::vector<unsigned char> out;
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
        if (c == '=') break;
        int v = val(c);
        if (v < 0) continue;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((unsigned char)((acc >> bits) & 0xFF));
        }
    }
    return out;
}

std::vector<std::array<double, 3>> palette_from_blob(const std::string& s) {
    auto bin = b64url_decode(s);
    if (bin.size() < 8) throw std::runtime_error("palette blob too short");
    uint64_t len = 0;
    std::memcpy(&len, bin.data(), 8);
    if (len % (3 * sizeof(double)) != 0 || len > (1u << 24)) throw std::runtime_error("bad palette blob size");
    std::vector<unsigned char> raw(len);
    uLongf out_len = uLongf(len);
    if (uncompress(raw.data(), &out_len, bin.data() + 8, uLong(bin.size() - 8)) != Z_OK || out_len != len)
        throw std::runtime_error("palette blob does not inflate");
    std::vector<std::array<double, 3>> pal(len / (3 * sizeof(double)));
    std::memcpy(pal.data(), raw.data(), len);
    return pal;
}

}  // namespace

flame flame_from_json(const json::value& js, const variation_db& db) {
    if (!js.is_object()) throw std::runtime_error("flame must be a JSON object");
    flame f;
    ANIMA_FIELD(f, js, center_x);
    ANIMA_FIELD(f, js, center_y);
    ANIMA_FIELD(f, js, scale);
    ANIMA_FIELD(f, js, rotate);
    ANIMA_FIELD(f, js, gamma);
    ANIMA_FIELD(f, js, brightness);
    ANIMA_FIELD(f, js, vibrancy);
    ANIMA_FIELD(f, js, highlight_power);
    ANIMA_FIELD(f, js, gamma_threshold);
    ANIMA_FIELD(f, js, mod_hue);
    ANIMA_FIELD(f, js, mod_sat);
    ANIMA_FIELD(f, js, mod_val);
    if (js.contains("name") && js.at("name").is_string()) f.name = js.at("name").str;
    if (!js.contains("xforms") || !js.at("xforms").is_array()) throw std::runtime_error("flame needs xforms");
    for (auto& x : *js.at("xforms").arr) f.xforms.push_back(xform_from(x, db));
    if (f.xforms.empty()) throw std::runtime_error("flame has no xforms");
    if (f.xforms.size() > 250) throw std::runtime_error("too many xforms");
    if (js.contains("final_xform")) f.final_xform = xform_from(js.at("final_xform"), db);
    if (js.contains("chaos")) {
        const auto& ch = js.at("chaos");
        if (!ch.is_array() || ch.arr->size() != f.xforms.size()) throw std::runtime_error("chaos must be n x n");
        std::vector<std::vector<anima>> table;
        for (auto& row : *ch.arr) {
            if (!row.is_array() || row.arr->size() != f.xforms.size()) throw std::runtime_error("chaos must be n x n");
            std::vector<anima> r;
            for (auto& v : *row.arr) r.push_back(anima_from(v));
            table.push_back(std::move(r));
        }
        f.chaos = std::move(table);
    }
    if (!js.contains("palette")) throw std::runtime_error("flame needs a palette");
    const auto& pal = js.at("palette");
    if (pal.is_string()) {
        f.palette = palette_from_blob(pal.str);
    } else if (pal.is_array()) {
        for (auto& e : *pal.arr) {
            if (!e.is_array() || e.arr->size() != 3) throw std::runtime_error("palette entries must be [h,s,v]");
            f.palette.push_back({(*e.arr)[0].as_number(), (*e.arr)[1].as_number(), (*e.arr)[2].as_number()});
        }
    } else {
        throw std::runtime_error("palette must be a string or an array");
    }
    if (f.palette.size() != 256) throw std::runtime_error("palette must have 256 entries");
    return f;
}

json::value flame_to_json(const flame& f) {
    auto o = json::value::make_object();
    auto& m = *o.obj;
    m["center_x"] = anima_to(f.center_x);
    m["center_y"] = anima_to(f.center_y);
    m["scale"] = anima_to(f.scale);
    m["rotate"] = anima_to(f.rotate);
    m["gamma"] = anima_to(f.gamma);
    m["brightness"] = anima_to(f.brightness);
    m["vibrancy"] = anima_to(f.vibrancy);
    m["highlight_power"] = anima_to(f.highlight_power);
    m["gamma_threshold"] = anima_to(f.gamma_threshold);
    m["mod_hue"] = anima_to(f.mod_hue);
    m["mod_sat"] = anima_to(f.mod_sat);
    m["mod_val"] = anima_to(f.mod_val);
    if (!f.name.empty()) m["name"] = json::value::make_string(f.name);
    auto xs = json::value::make_array();
    for (auto& x : f.xforms) xs.arr->push_back(xform_to(x));
    m["xforms"] = xs;
    if (f.final_xform) m["final_xform"] = xform_to(*f.final_xform);
    if (f.chaos) {
        auto ch = json::value::make_array();
        for (auto& row : *f.chaos) {
            auto r = json::value::make_array();
            for (auto& v : row) r.arr->push_back(anima_to(v));
            ch.arr->push_back(r);
        }
        m["chaos"] = ch;
    }
    auto pal = json::value::make_array();
    for (auto& [h, s, v] : f.palette) {
        auto e = json::value::make_array();
        e.arr->push_back(json::value::make_number(h));
        e.arr->push_back(json::value::make_number(s));
        e.arr->push_back(json::value::make_number(v));
        pal.arr->push_back(e);
    }
    m["palette"] = pal;
    return o;
}

// --------------------------------------------------------------- rotation

void fix_rotation(double* s, size_t sample_size, size_t flame_size, const size_t* indices, size_t n_indices,
                  int count) {
    constexpr double eps = 1e-10;
    // log(0) would make the cubic NaN; a zero column is given a tiny magnitude instead
    constexpr double log_floor = -700.0;
    for (size_t k = 0; k < n_indices; k++) {
        size_t idx = indices[k];
        for (int i = 0; i < count; i++) {
            size_t o = idx + size_t(i) * sample_size;
            double ax = std::atan2(s[o + 1], s[o]), mx = std::sqrt(s[o] * s[o] + s[o + 1] * s[o + 1]);
            double ay = std::atan2(s[o + 3], s[o + 2]), my = std::sqrt(s[o + 2] * s[o + 2] + s[o + 3] * s[o + 3]);
            double tx = s[o + 4], ty = s[o + 5];
            bool zx = mx == 0.0, zy = my == 0.0;
            if (zx && !zy) ax = ay;
            if (zy && !zx) ay = ax;
            s[o] = ax;
            s[o + 1] = ay;
            s[o + 2] = zx ? log_floor : std::log(mx);
            s[o + 3] = zy ? log_floor : std::log(my);
            s[o + 4] = tx;
            s[o + 5] = ty;
        }
        for (int i = 1; i < count; i++) {
            for (int j = 0; j < 2; j++) {
                size_t me = idx + size_t(i) * sample_size + size_t(j);
                size_t prev = idx + size_t(i - 1) * sample_size + size_t(j);
                double diff = s[me] - s[prev];
                if (diff > kPi + eps) s[me] -= 2 * kPi;
                else if (diff < -(kPi - eps)) s[me] += 2 * kPi;
            }
        }
    }
    for (size_t i = flame_size; i < sample_size; i += 3) {
        for (int j = 1; j < count; j++) {
            size_t me = i + size_t(j) * sample_size, prev = i + size_t(j - 1) * sample_size;
            double diff = s[me] - s[prev];
            if (diff > 180.0 + eps) s[me] -= 360.0;
            else if (diff < -(180.0 - eps)) s[me] += 360.0;
        }
    }
}

std::array<double, 3> rgb_to_hsv(double r, double g, double b) {
    double cmax = std::max({r, g, b}), cmin = std::min({r, g, b}), delta = cmax - cmin;
    std::array<double, 3> out{0.0, 0.0, cmax};
    if (delta > 0.0) {
        out[1] = delta / cmax;
        if (r == cmax) out[0] = 60.0 * std::fmod((g - b) / delta, 6.0);
        else if (g == cmax) out[0] = 60.0 * (((b - r) / delta) + 2.0);
        else out[0] = 60.0 * (((r - g) / delta) + 4.0);
    }
    return out;
}

std::vector<flame> align_by_weight(const std::vector<flame>& in) {
    std::vector<flame> out;
    bool any_final = false;
    for (auto& f : in) any_final |= f.final_xform.has_value();
    for (size_t k = 0; k < in.size(); k++) {
        flame f = in[k];
        f.xforms.clear();
        for (size_t j = 0; j < in.size(); j++) {
            for (auto xf : in[j].xforms) {
                if (j != k) xf.weight = 0.0;
                f.xforms.push_back(std::move(xf));
            }
        }
        f.chaos.reset();
        out.push_back(std::move(f));
    }
    if (any_final) {
        // every keyframe needs the same final-xform structure: pad with an
        // identity linear xform where a keyframe has none
        for (auto& f : out) {
            if (!f.final_xform) {
                xform id;
                id.weight = 0.0;
                id.color = 0.0;
                id.color_speed = 0.0;
                id.opacity = 1.0;
                vlink vl;
                vardata lin;
                lin.weight = 1.0;
                vl.variations["linear"] = lin;
                id.vchain.push_back(vl);
                f.final_xform = id;
            }
        }
    }
    return out;
}

}  // namespace rfk
