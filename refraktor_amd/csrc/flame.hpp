// Host-side flame data model and the packed parameter layout.
//
// The layout is the one librefrakt's device code indexes
// (src/librefrakt/include/librefrakt/flame_types.h:602-712):
//   [screen affine(6), plane affine(6), weight_sum(1),
//    chaos rows n x (weight_sum, w_0..w_{n-1}) when xaos is on,
//    xform_i { weight, color, color_speed, opacity,
//              vlink_j { affine a,d,b,e,c,f,
//                        variation_k { weight, user params (lexicographic),
//                                      precalc slots (lexicographic) } } },
//    final xform (same shape)]
// followed, in each packed sample, by 256 palette entries (h, s, v).
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <optional>
#include <string>
#include <variant>
#include <vector>

#include "json.hpp"
#include "vardb.hpp"

namespace rfk {

// A parameter value with an optional animation call (flame_types.h:47-100).
struct anima {
    using arg_t = std::variant<int, double, bool, std::string>;
    double t0 = 0.0;
    std::optional<std::string> call;
    std::map<std::string, arg_t> args;

    anima() = default;
    anima(double v) : t0(v) {}
    double sample(double t) const;
};

struct affine {
    anima a = 0.0, d = 0.0, b = 0.0, e = 0.0, c = 0.0, f = 0.0;
    static affine identity() { affine r; r.a = 1.0; r.e = 1.0; return r; }
};

// plain (already sampled) affine with librefrakt's scaled/rotated/translated
struct affine_v {
    double a, d, b, e, c, f;
    affine_v scaled(double s) const { return {a * s, d * s, b * s, e * s, c, f}; }
    affine_v rotated(double deg) const;
    affine_v translated(double x, double y) const { return {a, d, b, e, a * x + b * y + c, d * x + e * y + f}; }
};

struct vardata {
    anima weight = 0.0;
    std::map<std::string, anima> params;  // user parameters, lexicographic
    size_t precalc_count = 0;
    size_t size_reals() const { return 1 + params.size() + precalc_count; }
};

struct vlink {
    affine transform = affine::identity();
    anima mod_x = 0.0, mod_y = 0.0, mod_scale = 1.0, mod_rotate = 0.0;
    std::map<std::string, vardata> variations;  // lexicographic
    size_t size_reals() const {
        size_t s = 6;
        for (auto& [_, v] : variations) s += v.size_reals();
        return s;
    }
};

struct xform {
    anima weight = 0.0, color = 0.0, color_speed = 0.0, opacity = 1.0;
    std::vector<vlink> vchain;
    size_t size_reals() const {
        size_t s = 4;
        for (auto& v : vchain) s += v.size_reals();
        return s;
    }
};

struct flame {
    std::string name;
    std::vector<xform> xforms;
    std::optional<xform> final_xform;
    std::optional<std::vector<std::vector<anima>>> chaos;
    anima center_x = 0.0, center_y = 0.0, scale = 1.0, rotate = 0.0;
    anima gamma = 4.0, brightness = 4.0, vibrancy = 1.0, highlight_power = 0.0, gamma_threshold = 0.0;
    std::vector<std::array<double, 3>> palette;  // (h degrees, s, v)
    anima mod_hue = 0.0, mod_sat = 0.0, mod_val = 0.0;

    size_t size_reals() const;
    std::vector<size_t> affine_indices() const;
    affine_v screen_space(int w, int h, double t) const;
    affine_v plane_space(int w, int h, double t) const;
    void pack_sample(std::vector<double>& out, double t, int w, int h) const;
};

// Offsets of every named slot in the packed layout, for code generation.
struct var_slot {
    std::string name;
    int weight = -1;
    std::vector<std::pair<std::string, int>> user;     // user params
    std::vector<std::pair<std::string, int>> precalc;  // precalc params
};
struct vlink_slot {
    int aff = -1;
    std::vector<var_slot> vars;
};
struct xform_slot {
    int weight = -1, color = -1, color_speed = -1, opacity = -1;
    std::vector<vlink_slot> vlinks;
};
struct flame_layout {
    int size_reals = 0;
    int weight_sum = 12;
    int chaos_base = -1;  // -1 without xaos
    std::vector<xform_slot> xforms;
    std::optional<xform_slot> final_xform;
    std::vector<size_t> affine_indices;
    // reals per block on the device: the layout, then (without xaos) the
    // xform selection table that the per-sample precalc fills: n slots from
    // size_reals on and a flag, none of them part of the reference's layout
    int device_reals() const { return size_reals + (chaos_base < 0 ? int(xforms.size()) + 1 : 0); }
};

flame_layout make_layout(const flame& f, const variation_db& db);

// JSON (flame_types.cpp:247-485 document shape). Throws std::runtime_error.
flame flame_from_json(const json::value& js, const variation_db& db);
json::value flame_to_json(const flame& f);
// flam3 XML import (flame_types.cpp:131-239). Throws std::runtime_error.
flame flame_from_flam3(const std::string& xml, const variation_db& db);

// vardata with database defaults (flame_info.cpp:224-237)
vardata make_vardata(const variation_db& db, const std::string& name);

// Host polar conversion + angle unwrapping over `count` packed samples
// (flame_compiler.cpp:957-1035, generalised from 4 to `count` samples).
void fix_rotation(double* samples, size_t sample_size, size_t flame_size, const size_t* indices, size_t n_indices,
                  int count);

// librefrakt color::rgb_to_hsv (src/librefrakt/src/util/color.cpp:5-27)
std::array<double, 3> rgb_to_hsv(double r, double g, double b);

// Keyframe alignment by zero-weight concatenation (refraktor.cpp:146-198).
std::vector<flame> align_by_weight(const std::vector<flame>& in);

}  // namespace rfk
