#include "compiler.hpp"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "flang.hpp"

namespace rfk {

namespace {
#include "device_sources.inc"  // generated at build time: kFlameDeviceSrc, kFlameKernelsSrc
}

const std::string& device_runtime_source() {
    static const std::string s(kFlameDeviceSrc);
    return s;
}
const std::string& device_kernels_source() {
    static const std::string s(kFlameKernelsSrc);
    return s;
}

namespace {

std::string P(int i) { return "P[" + std::to_string(i) + "]"; }

// one vlink: pre-affine, zeroed accumulator, required common values, then
// every variation's code (flame.tpl:46-57, 71-79)
std::string emit_vlink(const variation_db& db, const vlink& vl, const vlink_slot& vs, int ind) {
    std::string pad(size_t(ind) * 4, ' ');
    std::ostringstream os;
    os << pad << "{\n";
    os << pad << "    const v2 p = fl::affine(P + " << vs.aff << ", q);\n";
    os << pad << "    v2 res = {Real(0), Real(0)};\n";
    // union of commons needed by the vlink's variations, dependencies first
    std::vector<std::string> commons;
    for (auto& v : vs.vars) {
        const variation_def* def = db.find(v.name);
        for (auto& c : db.required_common(*def)) {
            bool seen = false;
            for (auto& e : commons) seen |= e == c;
            if (!seen) commons.push_back(c);
        }
    }
    flang::emit_env cenv;
    cenv.aff_base = vs.aff;
    for (auto& c : commons) {
        const common_def* cd = db.find_common(c);
        os << pad << "    const Real common_" << c << " = " << flang::emit_expression(cd->expr.get(), cenv) << ";\n";
    }
    (void)vl;
    for (auto& v : vs.vars) {
        const variation_def* def = db.find(v.name);
        flang::emit_env env;
        env.weight = P(v.weight);
        env.aff_base = vs.aff;
        for (auto& [pn, off] : v.user) env.params[pn] = P(off);
        for (auto& [pn, off] : v.precalc) env.params[pn] = P(off);
        os << pad << "    // " << v.name << "\n";
        os << flang::emit_statements(def->apply.get(), env, ind + 1);
    }
    os << pad << "    q = res;\n";
    os << pad << "}\n";
    return os.str();
}

std::string emit_xform_fn(const variation_db& db, const xform& xf, const xform_slot& xs, const std::string& name) {
    std::ostringstream os;
    os << "RFK_INLINE void " << name << "(param_ptr P, v2& q, rng_t& rs) {\n";
    os << "    (void)rs;\n";
    for (size_t j = 0; j < xs.vlinks.size(); j++) {
        os << "    // vlink " << j << "\n";
        os << emit_vlink(db, xf.vchain[j], xs.vlinks[j], 1);
    }
    os << "}\n\n";
    return os.str();
}

std::string emit_precalc_vars(const variation_db& db, const xform_slot& xs) {
    std::ostringstream os;
    for (auto& vs : xs.vlinks) {
        for (auto& v : vs.vars) {
            const variation_def* def = db.find(v.name);
            if (!def->precalc) continue;
            flang::emit_env env;
            env.weight = P(v.weight);
            env.aff_base = vs.aff;
            for (auto& [pn, off] : v.user) env.params[pn] = P(off);
            for (auto& [pn, off] : v.precalc) env.params[pn] = P(off);
            os << "    // " << v.name << " precalc\n";
            os << flang::emit_statements(def->precalc.get(), env, 1);
        }
    }
    return os.str();
}

}  // namespace

std::string make_flame_functions(const variation_db& db, const flame& f, const flame_layout& L) {
    std::ostringstream os;
    const size_t n = f.xforms.size();

    for (size_t i = 0; i < n; i++) os << emit_xform_fn(db, f.xforms[i], L.xforms[i], "rfk_xform_" + std::to_string(i));
    if (f.final_xform) os << emit_xform_fn(db, *f.final_xform, *L.final_xform, "rfk_xform_final");

    // dispatch (flame.tpl:146-160): xform, then colour blend, returns opacity
    os << "RFK_INLINE Real rfk_dispatch(u32 xid, param_ptr P, v3& pt, rng_t& rs) {\n";
    os << "    v2 q = {pt.x, pt.y};\n";
    os << "    switch (xid) {\n";
    auto emit_case = [&](size_t idx, const xform_slot& xs, const std::string& fn) {
        os << "        case " << idx << ": {\n";
        os << "            " << fn << "(P, q, rs);\n";
        os << "            pt.x = q.x;\n            pt.y = q.y;\n";
        os << "            pt.z = pt.z * (Real(1) - " << P(xs.color_speed) << ") + " << P(xs.color) << " * "
           << P(xs.color_speed) << ";\n";
        os << "            return " << P(xs.opacity) << ";\n";
        os << "        }\n";
    };
    for (size_t i = 0; i < n; i++) emit_case(i, L.xforms[i], "rfk_xform_" + std::to_string(i));
    if (f.final_xform) emit_case(n, *L.final_xform, "rfk_xform_final");
    os << "        default: return Real(0);\n    }\n}\n\n";

    // cumulative-weight selection (flame.tpl:110-123)
    os << "RFK_INLINE u32 rfk_select_xform(param_ptr P, Real ratio) {\n";
    os << "    RFK_EXACT_FP\n";
    // IEEE-exact products and sums so the choice is reproducible off the GPU
    os << "    ratio = RFK_MUL_RN(ratio, " << P(L.weight_sum) << ");\n";
    if (L.chaos_base < 0 && n > 0) {
        // the same choice from the sample's table (rfk_precalc_sample): the
        // running sums' bits order like the sums themselves where the flag
        // says every weight and their total are finite and not negative, so
        // the wave-uniform choice is scalar loads and integer compares
        const int base = L.size_reals;
        os << "#ifndef RFK_DOUBLE\n";
        os << "    {\n";
        os << "        typedef const __attribute__((address_space(4))) int* table_ptr;\n";
        os << "        const table_ptr S = (table_ptr)(P + " << base << ");\n";
        // every slot read before the first compare: one wide scalar load
        os << "        int s[" << (n + 1) << "];\n";
        os << "        for (int i = 0; i < " << (n + 1) << "; i++) s[i] = S[i];\n";
        os << "        if (s[" << n << "] != 0) {\n";
        os << "            const int rb = __builtin_amdgcn_readfirstlane(__float_as_int(ratio));\n";
        for (size_t i = 0; i + 1 < n; i++) os << "            if (s[" << i << "] >= rb) return " << i << "u;\n";
        os << "            return u32(s[" << (n - 1) << "]);\n";
        os << "        }\n";
        os << "    }\n";
        os << "#endif\n";
    }
    os << "    Real rsum = Real(0);\n    u32 last_nonzero = 0;\n";
    for (size_t i = 0; i < n; i++) {
        const std::string w = P(L.xforms[i].weight);
        if (i + 1 < n) {
            os << "    if (" << w << " != Real(0) && RFK_ADD_RN(rsum, " << w << ") >= ratio) return " << i << "u;\n";
            os << "    rsum = RFK_ADD_RN(rsum, " << w << ");\n";
            os << "    if (" << w << " != Real(0)) last_nonzero = " << i << "u;\n";
        } else {
            os << "    return (" << w << " != Real(0)) ? " << i << "u : last_nonzero;\n";
        }
    }
    os << "}\n\n";

    if (L.chaos_base >= 0) {
        // xaos selection (flame.tpl:127-142): row `last` of the chaos table
        os << "RFK_INLINE u32 rfk_select_xform_chaos(param_ptr P, u32 last, Real ratio) {\n";
        os << "    RFK_EXACT_FP\n";
        os << "    if (last >= " << n << "u) return rfk_select_xform(P, ratio);\n";
        os << "    param_ptr row = P + " << L.chaos_base << " + last * " << (n + 1) << "u;\n";
        os << "    ratio = RFK_MUL_RN(ratio, row[0]);\n    Real rsum = Real(0);\n    u32 last_nonzero = 0;\n";
        for (size_t i = 0; i < n; i++) {
            const std::string w = "row[" + std::to_string(i + 1) + "]";
            if (i + 1 < n) {
                os << "    if (" << w << " != Real(0) && RFK_ADD_RN(rsum, " << w << ") >= ratio) return " << i << "u;\n";
                os << "    rsum = RFK_ADD_RN(rsum, " << w << ");\n";
                os << "    if (" << w << " != Real(0)) last_nonzero = " << i << "u;\n";
            } else {
                os << "    return (" << w << " != Real(0)) ? " << i << "u : last_nonzero;\n";
            }
        }
        os << "}\n\n";
    }

    // per-sample precalc (flame.tpl:166-187): weight sum, xaos rows, AA
    // jitter of the screen transform, then each variation's precalc block
    os << "RFK_INLINE void rfk_precalc_sample(Real* __restrict__ P, rng_t& rs) {\n";
    os << "    RFK_EXACT_FP\n";
    os << "    " << P(L.weight_sum) << " = ";
    for (size_t i = 0; i < n; i++) os << (i ? " + " : "") << P(L.xforms[i].weight);
    os << ";\n";
    if (L.chaos_base < 0 && n > 0) {
        // rfk_select_xform's table: slot i < n-1 the bits of the running sum
        // through weight i (-1 for a zero weight, which is never chosen),
        // slot n-1 the choice when no sum reaches the ratio, slot n the flag
        const int base = L.size_reals;
        os << "#ifndef RFK_DOUBLE\n    {\n";
        os << "        int* S = (int*)(P + " << base << ");\n";
        os << "        Real rsum = Real(0);\n        int last_nonzero = 0;\n        bool ok = true;\n";
        for (size_t i = 0; i < n; i++) {
            const std::string w = P(L.xforms[i].weight);
            // finite and not negative (-0 included), read from the bits
            os << "        { const u32 b = __float_as_uint(" << w << "); ok = ok && (b < 0x7F800000u || b == 0x80000000u); }\n";
            if (i + 1 < n) {
                os << "        rsum = RFK_ADD_RN(rsum, " << w << ");\n";
                os << "        S[" << i << "] = " << w << " != Real(0) ? __float_as_int(rsum) : -1;\n";
                os << "        if (" << w << " != Real(0)) last_nonzero = " << i << ";\n";
            } else {
                os << "        S[" << i << "] = " << w << " != Real(0) ? " << i << " : last_nonzero;\n";
            }
        }
        os << "        { const u32 b = __float_as_uint(" << P(L.weight_sum) << "); ok = ok && b < 0x7F800000u; }\n";
        os << "        S[" << n << "] = ok ? 1 : 0;\n";
        os << "    }\n#endif\n";
    }
    if (L.chaos_base >= 0) {
        for (size_t i = 0; i < n; i++) {
            int row = L.chaos_base + int(i * (n + 1));
            for (size_t j = 0; j < n; j++) {
                os << "    " << P(row + 1 + int(j)) << " *= " << P(L.xforms[j].weight) << ";\n";
                os << "    " << P(row) << " += " << P(row + 1 + int(j)) << ";\n";
            }
        }
    }
    os << "    {\n        v2 jitter = rs.randgauss(Real(1.0f / 3.0f));\n";
    os << "        " << P(4) << " += jitter.x;\n        " << P(5) << " += jitter.y;\n    }\n";
    for (size_t i = 0; i < n; i++) os << emit_precalc_vars(db, L.xforms[i]);
    if (f.final_xform) os << emit_precalc_vars(db, *L.final_xform);
    os << "}\n";
    return os.str();
}

std::string make_source(const variation_db& db, const flame& f, const flame_layout& L, const compile_options& o) {
    std::ostringstream os;
    uint32_t bits = 0;
    while ((1u << bits) < o.tpb) bits++;
    os << "// generated by refraktor_amd for a flame of " << f.xforms.size() << " xforms"
       << (f.final_xform ? " + final" : "") << ", " << L.size_reals << " reals\n";
    if (o.precision == 1) os << "#define RFK_DOUBLE 1\n";
    os << "#define RFK_TPB " << o.tpb << "\n";
    os << "#define RFK_TPB_BITS " << bits << "\n";
    os << "#define RFK_NUM_XFORMS " << f.xforms.size() << "\n";
    os << "#define RFK_HAS_FINAL " << (f.final_xform ? 1 : 0) << "\n";
    os << "#define RFK_USE_CHAOS " << (L.chaos_base >= 0 ? 1 : 0) << "\n";
    os << "#define RFK_SIZE_REALS " << L.size_reals << "\n";
    os << "#define RFK_ATOMIC64 " << (o.atomic_width == 64 ? 1 : 0) << "\n";
    os << "#define RFK_BIN_TILE_LOG2 " << o.bin_tile_log2 << "\n";
    os << "#define RFK_HIST64 " << (o.hist64 ? 1 : 0) << "\n";
    os << "#define RFK_BIN_QUEUE " << (o.bin_queue ? 1 : 0) << "\n";
    os << device_runtime_source() << "\n";
    os << "// ---------------------------------------------------------------- flame\n";
    os << make_flame_functions(db, f, L) << "\n";
    os << device_kernels_source() << "\n";
    return os.str();
}

std::vector<std::string> default_compile_flags() {
    // fast math, but NaN-aware and with contraction that the RFK_EXACT_FP
    // regions (selection, precalc sums, binning) can switch off
    return {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffast-math", "-fno-finite-math-only",
            "-ffp-contract=fast-honor-pragmas",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

namespace {

std::string cache_dir() {
    const char* env = std::getenv("RFK_CACHE_DIR");
    std::string d;
    if (env && *env) d = env;
    else if (const char* home = std::getenv("HOME"); home && *home) d = std::string(home) + "/.cache/refraktor_amd";
    else d = "/tmp/refraktor_amd_cache";
    // create parents
    std::string cur;
    for (size_t i = 0; i < d.size(); i++) {
        cur.push_back(d[i]);
        if (d[i] == '/' || i + 1 == d.size()) mkdir(cur.c_str(), 0755);
    }
    if (access(d.c_str(), W_OK) != 0) {
        d = "/tmp/refraktor_amd_cache";
        mkdir(d.c_str(), 0755);
    }
    return d;
}

}  // namespace

compile_result compile_hip(const std::string& source, const std::string& name, const std::vector<std::string>& options) {
    compile_result r;
    auto t0 = std::chrono::steady_clock::now();
    int hv_major = 0, hv_minor = 0;
    hiprtcVersion(&hv_major, &hv_minor);
    uint64_t h = fnv1a64(source);
    for (auto& o : options) h = fnv1a64(o, h);
    h = fnv1a64(std::to_string(hv_major) + "." + std::to_string(hv_minor) + "/rfk-co-1", h);
    char hex[32];
    std::snprintf(hex, sizeof hex, "%016llx", (unsigned long long)h);
    const bool use_cache = !std::getenv("RFK_NO_CACHE");
    std::string path;
    if (use_cache) {
        path = cache_dir() + "/" + hex + ".co";
        std::ifstream in(path, std::ios::binary);
        if (in) {
            r.code.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
            if (!r.code.empty()) {
                r.ok = true;
                r.cache_hit = true;
                r.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                return r;
            }
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, source.c_str(), (name + ".hip").c_str(), 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        r.log = "hiprtcCreateProgram failed";
        return r;
    }
    std::vector<const char*> opts;
    for (auto& o : options) opts.push_back(o.c_str());
    hiprtcResult cr = hiprtcCompileProgram(prog, int(opts.size()), opts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        r.log.resize(log_size);
        hiprtcGetProgramLog(prog, r.log.data());
        while (!r.log.empty() && r.log.back() == '\0') r.log.pop_back();
    }
    if (cr == HIPRTC_SUCCESS) {
        size_t code_size = 0;
        hiprtcGetCodeSize(prog, &code_size);
        r.code.resize(code_size);
        hiprtcGetCode(prog, r.code.data());
        r.ok = code_size > 0;
    }
    hiprtcDestroyProgram(&prog);
    if (r.ok && use_cache) {
        std::string tmp = path + ".tmp" + std::to_string(getpid());
        std::ofstream out(tmp, std::ios::binary);
        out.write(r.code.data(), std::streamsize(r.code.size()));
        out.close();
        std::rename(tmp.c_str(), path.c_str());
    }
    r.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

}  // namespace rfk
