// Flame compiler: DSL -> gfx950 HIP source -> hiprtc code object.
// Replaces rfkt::flame_compiler (src/librefrakt/src/flame_compiler.cpp:448-656).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "flame.hpp"
#include "vardb.hpp"

namespace rfk {

struct compile_options {
    int precision = 0;        // 0 f32, 1 f64
    uint32_t tpb = 256;       // walkers per block (power of two, 64..1024)
    uint32_t atomic_width = 32;
    uint32_t bin_tile_log2 = 0;  // 0: no LDS bin cache, else 2^n cached bins per block
    bool hist64 = false;         // u64 RGBA bins instead of u32
    bool bin_queue = false;      // draws to a record queue (queue_bin.hip) instead of atomics
};

// Generated source: configuration macros, the device runtime, the per-flame
// xform/selection/precalc functions and the iterate kernels.
std::string make_source(const variation_db& db, const flame& f, const flame_layout& L, const compile_options& o);
// Only the per-flame part (for inspection and tests).
std::string make_flame_functions(const variation_db& db, const flame& f, const flame_layout& L);

struct compile_result {
    bool ok = false;
    std::vector<char> code;
    std::string log;
    double ms = 0.0;
    bool cache_hit = false;
};

// hiprtc compile for gfx950 with a code-object cache keyed by the source hash.
compile_result compile_hip(const std::string& source, const std::string& name, const std::vector<std::string>& options);
std::vector<std::string> default_compile_flags();

// the embedded device runtime sources
const std::string& device_runtime_source();
const std::string& device_kernels_source();

}  // namespace rfk
