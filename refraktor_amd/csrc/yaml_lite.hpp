// A small YAML reader covering the subset the variation database files use:
// block mappings, literal block scalars (`|`, `|-`), plain and quoted
// scalars, flow sequences `[..]`, flow mappings `{..}`, block sequences and
// `#` comments. It replaces yaml-cpp for rfkt::initialize
// (src/librefrakt/src/flame_info.cpp:78-79 loads the two files with it).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rfk::yaml {

struct node {
    enum class kind { null, scalar, sequence, mapping };
    kind k = kind::null;
    std::string scalar;
    std::vector<node> items;                          // sequence
    std::vector<std::pair<std::string, node>> pairs;  // mapping, in file order

    bool is_scalar() const { return k == kind::scalar; }
    bool is_sequence() const { return k == kind::sequence; }
    bool is_mapping() const { return k == kind::mapping; }
    const node* find(const std::string& key) const {
        if (!is_mapping()) return nullptr;
        for (auto& [kk, v] : pairs)
            if (kk == key) return &v;
        return nullptr;
    }
};

struct parse_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

node parse(const std::string& text);
node parse_file(const std::string& path);

}  // namespace rfk::yaml
