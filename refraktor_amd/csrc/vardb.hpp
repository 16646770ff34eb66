// Variation database: loads `variations_new.yml` + `common.yml` from a config
// directory, parses and type-checks every entry, and keeps the ASTs for code
// generation. Mirrors rfkt::flamedb / rfkt::initialize
// (src/librefrakt/include/librefrakt/flame_info.h:18-159,
//  src/librefrakt/src/flame_info.cpp:60-237): variations are registered in
// sorted-name order, entries that fail to parse or validate are reported and
// skipped, parameters are kept in lexicographic order with `precalc`-tagged
// ones stored after the user parameters.
#pragma once

#include <cstdint>
#include <map>
#include <optional>
#include <set>
#include <string>
#include <vector>

#include "flang.hpp"

namespace rfk {

enum class pad_type { none, identity, rotated, affine };

struct param_def {
    std::string name;
    std::optional<double> default_value;
    std::optional<double> identity_value;
    std::set<std::string> tags;
    bool is_precalc() const { return tags.count("precalc") > 0; }
};

struct variation_def {
    std::string name;
    std::string source;
    std::optional<std::string> precalc_source;
    pad_type pad = pad_type::none;
    std::map<std::string, param_def> params;  // lexicographic
    flang::node_ptr apply;
    flang::node_ptr precalc;
    std::vector<std::string> common_used;     // direct common.X references of apply

    std::vector<std::string> user_params() const;     // non-precalc, lexicographic
    std::vector<std::string> precalc_params() const;  // precalc-tagged, lexicographic
};

struct common_def {
    std::string name;
    std::string source;
    flang::node_ptr expr;
    std::vector<std::string> deps;  // common values it references
};

class variation_db {
public:
    // Load both YAML files from `dir`. Returns false if a file is missing or
    // unreadable; individual bad entries only land in errors().
    bool load(const std::string& dir, std::string& err);

    bool add_common(const std::string& name, const std::string& source, std::string& err);
    bool add_variation(variation_def def, std::string& err);

    const variation_def* find(const std::string& name) const {
        auto it = variations_.find(name);
        return it == variations_.end() ? nullptr : &it->second;
    }
    const common_def* find_common(const std::string& name) const {
        auto it = common_.find(name);
        return it == common_.end() ? nullptr : &it->second;
    }
    const std::map<std::string, variation_def>& variations() const { return variations_; }
    const std::map<std::string, common_def>& commons() const { return common_; }
    const std::vector<std::string>& errors() const { return errors_; }
    const std::set<std::string>& pad_set(pad_type t) const { return pad_sets_.at(int(t)); }
    uint64_t hash() const { return hash_; }

    // common values a variation needs, dependencies first (flame_compiler.cpp:428-446)
    std::vector<std::string> required_common(const variation_def& v) const;

private:
    void recalc_hash();
    std::map<std::string, variation_def> variations_;
    std::map<std::string, common_def> common_;
    std::vector<std::string> errors_;
    std::map<int, std::set<std::string>> pad_sets_{{0, {}}, {1, {}}, {2, {}}, {3, {}}};
    uint64_t hash_ = 0;
};

uint64_t fnv1a64(const void* data, size_t n, uint64_t h = 1469598103934665603ull);
inline uint64_t fnv1a64(const std::string& s, uint64_t h = 1469598103934665603ull) { return fnv1a64(s.data(), s.size(), h); }

}  // namespace rfk
