// Minimal JSON value, parser and writer used for the flame description that
// crosses the C ABI (the same document shape librefrakt's flame::serialize
// writes, flame_types.cpp:399-485).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rfk::json {

struct value;
using object = std::map<std::string, value>;
using array = std::vector<value>;

struct value {
    enum class kind { null, boolean, number, string, array, object };
    kind k = kind::null;
    bool b = false;
    double num = 0.0;
    bool is_integer = false;
    std::string str;
    std::shared_ptr<array> arr;
    std::shared_ptr<object> obj;

    value() = default;
    static value make_number(double v, bool integral = false) {
        value r; r.k = kind::number; r.num = v; r.is_integer = integral; return r;
    }
    static value make_string(std::string s) { value r; r.k = kind::string; r.str = std::move(s); return r; }
    static value make_bool(bool v) { value r; r.k = kind::boolean; r.b = v; return r; }
    static value make_array() { value r; r.k = kind::array; r.arr = std::make_shared<array>(); return r; }
    static value make_object() { value r; r.k = kind::object; r.obj = std::make_shared<object>(); return r; }

    bool is_null() const { return k == kind::null; }
    bool is_number() const { return k == kind::number; }
    bool is_string() const { return k == kind::string; }
    bool is_array() const { return k == kind::array; }
    bool is_object() const { return k == kind::object; }
    bool is_bool() const { return k == kind::boolean; }

    bool contains(const std::string& key) const { return is_object() && obj->count(key) > 0; }
    const value& at(const std::string& key) const {
        if (!is_object()) throw std::runtime_error("json: not an object");
        auto it = obj->find(key);
        if (it == obj->end()) throw std::runtime_error("json: missing key '" + key + "'");
        return it->second;
    }
    const array& elements() const {
        if (!is_array()) throw std::runtime_error("json: not an array");
        return *arr;
    }
    double as_number() const {
        if (!is_number()) throw std::runtime_error("json: not a number");
        return num;
    }
};

value parse(const std::string& text);
std::string dump(const value& v);

}  // namespace rfk::json
