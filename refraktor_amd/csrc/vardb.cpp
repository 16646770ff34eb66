#include "vardb.hpp"

#include <algorithm>
#include <cstdlib>
#include <functional>

#include "yaml_lite.hpp"

namespace rfk {

uint64_t fnv1a64(const void* data, size_t n, uint64_t h) {
    auto p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::vector<std::string> variation_def::user_params() const {
    std::vector<std::string> out;
    for (auto& [n, p] : params)
        if (!p.is_precalc()) out.push_back(n);
    return out;
}

std::vector<std::string> variation_def::precalc_params() const {
    std::vector<std::string> out;
    for (auto& [n, p] : params)
        if (p.is_precalc()) out.push_back(n);
    return out;
}

bool variation_db::add_common(const std::string& name, const std::string& source, std::string& err) {
    try {
        common_def c;
        c.name = name;
        c.source = source;
        c.expr = flang::parse_expression(source);
        flang::collect_common(c.expr.get(), c.deps);
        std::set<std::string> names;
        for (auto& [n, _] : common_) names.insert(n);
        for (auto& d : c.deps) names.insert(d);  // forward references resolved at codegen
        flang::check_common(c.expr.get(), names);
        common_[name] = std::move(c);
        recalc_hash();
        return true;
    } catch (const std::exception& e) {
        err = "cannot parse common `" + name + "`: " + e.what();
        errors_.push_back(err);
        return false;
    }
}

bool variation_db::add_variation(variation_def def, std::string& err) {
    try {
        def.apply = flang::parse_statements(def.source);
        if (def.precalc_source) def.precalc = flang::parse_statements(*def.precalc_source);
        flang::check_env env;
        for (auto& [n, _] : common_) env.common.insert(n);
        for (auto& [n, _] : def.params) env.params.insert(n);
        flang::check_variation(def.apply.get(), env);
        if (def.precalc) {
            env.precalc = true;
            flang::check_variation(def.precalc.get(), env);
        }
        def.common_used.clear();
        flang::collect_common(def.apply.get(), def.common_used);
    } catch (const std::exception& e) {
        err = "cannot validate variation `" + def.name + "`: " + e.what();
        errors_.push_back(err);
        return false;
    }
    if (auto it = variations_.find(def.name); it != variations_.end()) {
        pad_sets_[int(it->second.pad)].erase(def.name);
    }
    pad_sets_[int(def.pad)].insert(def.name);
    variations_[def.name] = std::move(def);
    recalc_hash();
    return true;
}

bool variation_db::load(const std::string& dir, std::string& err) {
    yaml::node vdefs, cdefs;
    try {
        vdefs = yaml::parse_file(dir + "/variations_new.yml");
        cdefs = yaml::parse_file(dir + "/common.yml");
    } catch (const std::exception& e) {
        err = e.what();
        return false;
    }
    if (!vdefs.is_mapping() || !cdefs.is_mapping()) {
        err = "variation files must be YAML mappings";
        return false;
    }
    std::string e2;
    for (auto& [name, v] : cdefs.pairs) {
        if (!v.is_scalar()) {
            errors_.push_back("common `" + name + "` is not a scalar");
            continue;
        }
        add_common(name, v.scalar, e2);
    }
    std::vector<std::string> names;
    for (auto& [name, _] : vdefs.pairs) names.push_back(name);
    std::sort(names.begin(), names.end());
    for (auto& name : names) {
        const yaml::node* y = vdefs.find(name);
        variation_def def;
        def.name = name;
        if (y->is_scalar()) {
            def.source = y->scalar;
        } else if (y->is_mapping()) {
            const yaml::node* src = y->find("src");
            if (!src || !src->is_scalar()) {
                errors_.push_back("variation `" + name + "` has no src");
                continue;
            }
            def.source = src->scalar;
            if (auto pc = y->find("precalc"); pc && pc->is_scalar()) def.precalc_source = pc->scalar;
            if (auto tags = y->find("tags"); tags && tags->is_sequence()) {
                for (auto& t : tags->items) {
                    if (t.scalar == "pad_affine") def.pad = pad_type::affine;
                    else if (t.scalar == "pad_rotated") def.pad = pad_type::rotated;
                    else if (t.scalar == "pad_identity") def.pad = pad_type::identity;
                }
            }
            if (auto params = y->find("param"); params && params->is_mapping()) {
                for (auto& [pname, pnode] : params->pairs) {
                    param_def pd;
                    pd.name = pname;
                    if (auto d = pnode.find("default"); d && d->is_scalar()) pd.default_value = std::strtod(d->scalar.c_str(), nullptr);
                    if (auto d = pnode.find("identity"); d && d->is_scalar()) pd.identity_value = std::strtod(d->scalar.c_str(), nullptr);
                    if (auto t = pnode.find("tags"); t && t->is_sequence())
                        for (auto& tag : t->items) pd.tags.insert(tag.scalar);
                    def.params[pname] = pd;
                }
            }
        } else {
            errors_.push_back("variation `" + name + "` is neither a scalar nor a mapping");
            continue;
        }
        add_variation(std::move(def), e2);
    }
    return true;
}

std::vector<std::string> variation_db::required_common(const variation_def& v) const {
    std::vector<std::string> out;
    std::function<void(const std::vector<std::string>&)> visit = [&](const std::vector<std::string>& names) {
        for (auto& n : names) {
            if (std::find(out.begin(), out.end(), n) != out.end()) continue;
            auto c = find_common(n);
            if (!c) continue;
            visit(c->deps);
            if (std::find(out.begin(), out.end(), n) == out.end()) out.push_back(n);
        }
    };
    visit(v.common_used);
    return out;
}

void variation_db::recalc_hash() {
    uint64_t h = 1469598103934665603ull;
    for (auto& [n, v] : variations_) {
        h = fnv1a64(n, h);
        h = fnv1a64(v.source, h);
        if (v.precalc_source) h = fnv1a64(*v.precalc_source, h);
        for (auto& [pn, p] : v.params) {
            h = fnv1a64(pn, h);
            h = fnv1a64(std::string(p.is_precalc() ? "P" : "U"), h);
        }
    }
    for (auto& [n, c] : common_) {
        h = fnv1a64(n, h);
        h = fnv1a64(c.source, h);
    }
    hash_ = h;
}

}  // namespace rfk
