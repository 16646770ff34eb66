// C ABI implementation: handles, launch plumbing and state ownership.
// Replaces rfkt::flame_kernel / saved_state / tonemapper host code
// (src/librefrakt/src/flame_compiler.cpp:866-1108,
//  src/librefrakt/src/image/tonemapper.cpp:37-56) with explicit HIP calls,
// status codes instead of exit(), and a pinned host stop flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/refrakt.h"
#include "compiler.hpp"
#include "device/flame_device.h"
#include "flame.hpp"
#include "json.hpp"
#include "static_kernels.hpp"
#include "vardb.hpp"

struct rfk_db {
    rfk::variation_db db;
};

struct rfk_flame {
    rfk::flame f;
};

struct rfk_kernel {
    rfk::flame f;
    rfk::flame_layout L;
    rfk::compile_options opt;
    hipModule_t mod = nullptr;
    hipFunction_t f_iterate = nullptr;
    hipFunction_t f_precalc = nullptr;
    hipFunction_t f_batch = nullptr;
    double compile_ms = 0.0;
    bool cache_hit = false;
    int device = 0;
};

struct rfk_state {
    rfk_kernel* k = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t width = 0, height = 0, walkers = 0, tpb = 0, blocks = 0;
    uint32_t size_reals = 0, param_stride = 0, seed = 0;
    int precision = 0;
    void* params = nullptr;
    uint32_t* palettes = nullptr;
    float *wx = nullptr, *wy = nullptr, *wc = nullptr;
    uint32_t* wlast = nullptr;
    uint32_t* block_pass = nullptr;
    uint32_t* hist = nullptr;
    unsigned long long* counters = nullptr;        // device [8]
    unsigned long long* counters_host = nullptr;   // pinned [8 + 98]
    unsigned long long* dstats = nullptr;          // device [97 + 1], = counters + 8 (one allocation)
    uint32_t* stop_host = nullptr;                 // pinned, mapped
    uint32_t* stop_dev = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_result = nullptr;  // a bin call's counters have reached the host
    double quality = 0.0;
    uint64_t passes_done = 0;
    bool pending = false;
    uint64_t pending_passes = 0;
    double wallclock_khz = 100000.0;
    uint32_t stream_id = 0;
    bool hist64 = false;
    size_t hist_words() const { return size_t(width) * height * (hist64 ? 8u : 4u); }
    // warmup scratch: keyframe samples, their segments, affine kind table
    double *d_samples = nullptr, *d_segments = nullptr;
    int8_t* d_kind = nullptr;
    uint32_t scratch_samples = 0;
    double* staging = nullptr;  // pinned host copy of the samples being uploaded
    size_t staging_reals = 0;
    hipEvent_t ev_staging = nullptr;
    // queue binning scratch (kernels compiled with bin_queue): record
    // buffers for chunks of at most chunk_passes passes, allocated for
    // queue_passes passes (grown on demand by ensure_queue)
    unsigned long long *queue = nullptr, *sorted = nullptr;
    uint32_t *qcounts = nullptr, *qnseg = nullptr, *qtiles = nullptr;
    void* qsegments = nullptr;
    uint32_t chunk_passes = 0, queue_passes = 0, max_segments = 0;
    // density estimation: per-level 1/Z table for the last (max, min) radius
    // pair, and the per-pixel level scratch
    float* de_inv_z = nullptr;
    int* de_levels = nullptr;
    float de_max_radius = -1.0f, de_min_radius = -1.0f;
};

struct rfk_batch {
    rfk_kernel* k = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t frames = 0, width = 0, height = 0, tpb = 0, blocks_per_frame = 0, total_blocks = 0;
    uint32_t param_stride = 0, seed = 0, fuse = 0, grid = 0;
    void* params = nullptr;
    uint32_t* palettes = nullptr;
    uint32_t* hist = nullptr;
    uint32_t *d_item_frame = nullptr, *d_item_block = nullptr, *d_frame_row0 = nullptr, *d_frame_blocks = nullptr,
             *d_frame_passes = nullptr, *d_queue = nullptr, *d_row_stream = nullptr, *d_row_local = nullptr;
    unsigned long long* d_frame_counters = nullptr;
    unsigned long long* h_frame_counters = nullptr;
    std::vector<uint32_t> frame_passes;
    uint32_t nitems = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    bool hist64 = false;
    // queue binning: launches of chunks, a chunk = up to slots_per_chunk slots,
    // a slot = one frame's next slice of at most slice_passes passes
    struct chunk {
        uint32_t first, nslots, passes;
    };
    bool queue = false;
    std::vector<chunk> chunks;
    uint32_t slots_per_chunk = 0, slice_passes = 0, tiles_per_frame = 0, max_segments = 0;
    uint32_t *d_slots = nullptr, *d_slot_frame = nullptr, *qcounts = nullptr, *qtiles = nullptr, *qnseg = nullptr;
    float *wx = nullptr, *wy = nullptr, *wc = nullptr;
    void *records = nullptr, *sorted = nullptr, *qsegments = nullptr;
    size_t frame_words() const { return size_t(width) * height * (hist64 ? 8u : 4u); }
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

namespace rfk {
// shared by every translation unit of the library (rfk_last_error reports it)
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace rfk

namespace {

#define RFK_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(RFK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(RFK_ERR_NODEVICE, "no HIP device visible");
    return RFK_OK;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) + 16);
}

void free_state_buffers(rfk_state* s) {
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(s->params); f(s->palettes); f(s->wx); f(s->wy); f(s->wc); f(s->wlast); f(s->block_pass); f(s->hist);
    f(s->queue); f(s->sorted); f(s->qcounts); f(s->qnseg); f(s->qsegments); f(s->qtiles);
    f(s->d_samples); f(s->d_segments); f(s->d_kind); f(s->de_inv_z); f(s->de_levels);
    if (s->ev_staging) (void)hipEventSynchronize(s->ev_staging);
    if (s->staging) (void)hipHostFree(s->staging);
    if (s->ev_staging) (void)hipEventDestroy(s->ev_staging);
    f(s->counters);  // dstats lies inside it
    if (s->counters_host) (void)hipHostFree(s->counters_host);
    if (s->stop_host) (void)hipHostFree(s->stop_host);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->ev_result) (void)hipEventDestroy(s->ev_result);
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
}

rfk_iter_args base_args(rfk_state* s) {
    rfk_iter_args a;
    std::memset(&a, 0, sizeof a);
    a.params = static_cast<const float*>(s->params);
    a.palettes = s->palettes;
    a.wx = s->wx;
    a.wy = s->wy;
    a.wc = s->wc;
    a.wlast = s->wlast;
    a.block_pass = s->block_pass;
    a.hist = s->hist;
    a.stop_flag = s->stop_dev;
    a.counters = reinterpret_cast<u64*>(s->counters);
    a.width = s->width;
    a.height = s->height;
    a.param_stride = s->param_stride;
    a.seed = s->seed;
    a.stream = s->stream_id;
    a.total_walkers = s->walkers;
    a.check_every = 16;
    a.nsamples = s->blocks;
    a.queue = reinterpret_cast<u64*>(s->queue);
    a.qcounts = s->qcounts;
    a.tiles_x = rfk::sk::queue_tiles_x(s->width);
    a.ntiles = rfk::sk::queue_tiles(s->width, s->height);
    return a;
}

int launch_iterate(rfk_state* s, rfk_iter_args& a) {
    void* argv[] = {&a};
    // queue binning kernels count records per screen tile in dynamic LDS
    const unsigned lds = s->k->opt.bin_queue ? a.ntiles * 4u : 0u;
    RFK_HIP(hipModuleLaunchKernel(s->k->f_iterate, s->blocks, 1, 1, s->tpb, 1, 1, lds, s->stream, argv, nullptr));
    return RFK_OK;
}

// Queue binning's record buffers (8 B per record, twice, and the work-item
// list) hold one chunk of min(chunk_passes, passes) passes: a state whose
// calls bin 2^25 draws keeps 2 x 256 MiB instead of 2 x 1 GiB. A call with a
// larger chunk regrows them once the state's earlier work has finished.
int ensure_queue(rfk_state* s, uint64_t passes) {
    const uint32_t need = uint32_t(std::min<uint64_t>(s->chunk_passes, passes));
    if (need <= s->queue_passes) return RFK_OK;
    RFK_HIP(hipStreamSynchronize(s->stream));
    for (void* p : {static_cast<void*>(s->queue), static_cast<void*>(s->sorted), s->qsegments})
        if (p) (void)hipFree(p);
    s->queue = s->sorted = nullptr;
    s->qsegments = nullptr;
    s->queue_passes = 0;
    const size_t q = size_t(need) * s->walkers;
    s->max_segments = rfk::sk::queue_max_segments(q, rfk::sk::queue_tiles(s->width, s->height));
    RFK_HIP(hipMalloc(&s->queue, q * 8));
    RFK_HIP(hipMalloc(&s->sorted, q * 8));
    RFK_HIP(hipMalloc(&s->qsegments, size_t(s->max_segments) * 16));
    s->queue_passes = need;
    return RFK_OK;
}

// Queue binning: chunks of at most `queue_passes` passes, each followed by
// the sort and per-tile sum of its records. With an iteration budget every
// chunk is enqueued at once (a bailout makes the remaining chunks record
// nothing); without one the host checks the device's stop count after every
// chunk.
int bin_queued(rfk_state* s, rfk_iter_args& a, uint64_t passes, bool bounded) {
    a.queue = reinterpret_cast<u64*>(s->queue);
    for (uint64_t done = 0; done < passes;) {
        const uint32_t n = uint32_t(std::min<uint64_t>(s->queue_passes, passes - done));
        a.npasses = n;
        if (int rc = launch_iterate(s, a)) return rc;
        RFK_HIP(rfk::sk::queue_bin(s->queue, size_t(n) * s->tpb, s->blocks, s->sorted, s->qcounts, s->qtiles,
                                   s->qsegments, s->max_segments, s->qnseg, s->width, s->height, s->hist, s->hist64,
                                   s->k->opt.atomic_width == 64, s->stream));
        done += n;
        if (!bounded && done < passes) {
            RFK_HIP(hipMemcpyAsync(s->counters_host, s->counters, 8 * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, s->stream));
            RFK_HIP(hipStreamSynchronize(s->stream));
            if (s->counters_host[RFK_CNT_STOPPED] != 0) break;
        }
    }
    return RFK_OK;
}

// zero the counters and the density statistics behind them, the running
// minimum start clock at ~0 (fill kernels: no host buffer is read after this
// returns)
int reset_counters(rfk_state* s) {
    uint32_t* c = reinterpret_cast<uint32_t*>(s->counters);
    RFK_HIP(rfk::sk::fill_u32_except(c, 0u, 2 * (8 + 98), 2 * RFK_CNT_START, 2, 0xFFFFFFFFu, s->stream));
    return RFK_OK;
}

}  // namespace

extern "C" {

const char* rfk_last_error(void) { return g_last_error.c_str(); }
const char* rfk_version(void) { return "refraktor_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------------- DB

int rfk_db_load(const char* config_dir, rfk_db** out, int* n_errors) {
    if (!config_dir || !out) return fail(RFK_ERR_INVALID, "null argument");
    auto db = std::make_unique<rfk_db>();
    std::string err;
    if (!db->db.load(config_dir, err)) return fail(RFK_ERR_PARSE, err);
    if (n_errors) *n_errors = int(db->db.errors().size());
    *out = db.release();
    return RFK_OK;
}

int rfk_db_free(rfk_db* db) {
    delete db;
    return RFK_OK;
}

int rfk_db_variation_count(const rfk_db* db) { return db ? int(db->db.variations().size()) : 0; }

const char* rfk_db_variation_name(const rfk_db* db, int i) {
    if (!db || i < 0 || i >= int(db->db.variations().size())) return nullptr;
    auto it = db->db.variations().begin();
    std::advance(it, i);
    return it->first.c_str();
}

const char* rfk_db_error(const rfk_db* db, int i) {
    if (!db || i < 0 || i >= int(db->db.errors().size())) return nullptr;
    return db->db.errors()[size_t(i)].c_str();
}

uint64_t rfk_db_hash(const rfk_db* db) { return db ? db->db.hash() : 0; }

// ------------------------------------------------------------------- flame

int rfk_flame_from_json(const rfk_db* db, const char* json, rfk_flame** out) {
    if (!db || !json || !out) return fail(RFK_ERR_INVALID, "null argument");
    try {
        auto js = rfk::json::parse(json);
        auto f = std::make_unique<rfk_flame>();
        f->f = rfk::flame_from_json(js, db->db);
        *out = f.release();
        return RFK_OK;
    } catch (const std::exception& e) {
        return fail(RFK_ERR_PARSE, e.what());
    }
}

int rfk_flame_from_flam3(const rfk_db* db, const char* xml, rfk_flame** out) {
    if (!db || !xml || !out) return fail(RFK_ERR_INVALID, "null argument");
    try {
        auto f = std::make_unique<rfk_flame>();
        f->f = rfk::flame_from_flam3(xml, db->db);
        *out = f.release();
        return RFK_OK;
    } catch (const std::exception& e) {
        return fail(RFK_ERR_PARSE, e.what());
    }
}

int rfk_flame_to_json(const rfk_flame* f, char* buf, size_t cap, size_t* needed) {
    if (!f) return fail(RFK_ERR_INVALID, "null flame");
    std::string s = rfk::json::dump(rfk::flame_to_json(f->f));
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return RFK_OK;
}

int rfk_flame_free(rfk_flame* f) {
    delete f;
    return RFK_OK;
}

int rfk_flame_size_reals(const rfk_flame* f, size_t* out) {
    if (!f || !out) return fail(RFK_ERR_INVALID, "null argument");
    *out = f->f.size_reals();
    return RFK_OK;
}

int rfk_flame_affine_indices(const rfk_flame* f, size_t* out, size_t cap, size_t* count) {
    if (!f) return fail(RFK_ERR_INVALID, "null flame");
    auto idx = f->f.affine_indices();
    if (count) *count = idx.size();
    for (size_t i = 0; out && i < idx.size() && i < cap; i++) out[i] = idx[i];
    return RFK_OK;
}

int rfk_flame_pack_samples(const rfk_flame* f, double start, double offset, int count, int w, int h, double* out,
                           size_t cap) {
    if (!f || !out || count <= 0 || w <= 0 || h <= 0) return fail(RFK_ERR_INVALID, "bad argument");
    std::vector<double> v;
    for (int i = 0; i < count; i++) f->f.pack_sample(v, start + offset * i, w, h);
    if (v.size() > cap) return fail(RFK_ERR_INVALID, "output buffer too small: need " + std::to_string(v.size()));
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return RFK_OK;
}

int rfk_flame_xform_count(const rfk_flame* f, int* n_xforms, int* has_final) {
    if (!f) return fail(RFK_ERR_INVALID, "null flame");
    if (n_xforms) *n_xforms = int(f->f.xforms.size());
    if (has_final) *has_final = f->f.final_xform ? 1 : 0;
    return RFK_OK;
}

int rfk_flames_align(const rfk_db* db, rfk_flame* const* in, int n, rfk_flame** out) {
    if (!db || !in || !out || n <= 0) return fail(RFK_ERR_INVALID, "bad argument");
    std::vector<rfk::flame> fl;
    for (int i = 0; i < n; i++) fl.push_back(in[i]->f);
    auto aligned = rfk::align_by_weight(fl);
    for (int i = 0; i < n; i++) {
        out[i] = new rfk_flame;
        out[i]->f = std::move(aligned[size_t(i)]);
    }
    return RFK_OK;
}

// ---------------------------------------------------------------- compiler

static int make_options(const rfk_exec_config* cfg, int precision, rfk::compile_options& o) {
    o.precision = precision;
    o.tpb = (cfg && cfg->threads_per_block) ? cfg->threads_per_block : 256;
    o.atomic_width = (cfg && cfg->atomic_width) ? cfg->atomic_width : 32;
    if (o.tpb != 64 && o.tpb != 128 && o.tpb != 256 && o.tpb != 512 && o.tpb != 1024)
        return fail(RFK_ERR_INVALID, "threads_per_block must be 64, 128, 256, 512 or 1024");
    if (o.atomic_width != 32 && o.atomic_width != 64) return fail(RFK_ERR_INVALID, "atomic_width must be 32 or 64");
    if (precision != RFK_PRECISION_F32 && precision != RFK_PRECISION_F64) return fail(RFK_ERR_INVALID, "bad precision");
    o.bin_tile_log2 = cfg ? cfg->bin_tile_log2 : 0;
    const uint32_t hb = (cfg && cfg->hist_bits) ? cfg->hist_bits : 32u;
    if (hb != 32 && hb != 64) return fail(RFK_ERR_INVALID, "hist_bits must be 32 or 64");
    o.hist64 = hb == 64;
    o.bin_queue = cfg && cfg->bin_queue;
    if (o.bin_queue && o.bin_tile_log2) return fail(RFK_ERR_INVALID, "bin_queue and bin_tile_log2 exclude each other");
    if (o.bin_tile_log2 != 0 && (o.bin_tile_log2 < 6 || o.bin_tile_log2 > 12))
        return fail(RFK_ERR_INVALID, "bin_tile_log2 must be 0 or 6..12");
    // static LDS of one workgroup: palette, two point buffers, bin cache
    const size_t lds = 1024 + 2 * 3 * size_t(o.tpb) * (precision ? 8 : 4) +
                       (o.bin_tile_log2 ? (size_t(20) << o.bin_tile_log2) : 0) + 64;
    if (lds > 160u * 1024u) return fail(RFK_ERR_INVALID, "launch shape needs more than 160 KiB of LDS per workgroup");
    return RFK_OK;
}

int rfk_make_source(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg, char* buf,
                    size_t cap, size_t* needed) {
    if (!db || !f) return fail(RFK_ERR_INVALID, "null argument");
    rfk::compile_options o;
    if (int rc = make_options(cfg, precision, o)) return rc;
    try {
        auto L = rfk::make_layout(f->f, db->db);
        std::string src = rfk::make_source(db->db, f->f, L, o);
        if (needed) *needed = src.size() + 1;
        if (buf && cap) {
            size_t n = std::min(cap - 1, src.size());
            std::memcpy(buf, src.data(), n);
            buf[n] = '\0';
        }
        return RFK_OK;
    } catch (const std::exception& e) {
        return fail(RFK_ERR_PARSE, e.what());
    }
}

int rfk_compile(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg, rfk_kernel** out,
                char* log, size_t log_cap) {
    if (!db || !f || !out) return fail(RFK_ERR_INVALID, "null argument");
    if (log && log_cap) log[0] = '\0';
    auto k = std::make_unique<rfk_kernel>();
    if (int rc = make_options(cfg, precision, k->opt)) return rc;
    std::string src;
    try {
        k->f = f->f;
        k->L = rfk::make_layout(k->f, db->db);
        src = rfk::make_source(db->db, k->f, k->L, k->opt);
    } catch (const std::exception& e) {
        return fail(RFK_ERR_PARSE, e.what());
    }
    auto cr = rfk::compile_hip(src, "rfk_flame", rfk::default_compile_flags());
    if (log && log_cap) {
        size_t n = std::min(log_cap - 1, cr.log.size());
        std::memcpy(log, cr.log.data(), n);
        log[n] = '\0';
    }
    if (!cr.ok) return fail(RFK_ERR_COMPILE, "hiprtc compile failed:\n" + cr.log);
    k->compile_ms = cr.ms;
    k->cache_hit = cr.cache_hit;
    if (int rc = check_device()) return rc;
    RFK_HIP(hipGetDevice(&k->device));
    RFK_HIP(hipModuleLoadData(&k->mod, cr.code.data()));
    RFK_HIP(hipModuleGetFunction(&k->f_iterate, k->mod, "rfk_iterate"));
    RFK_HIP(hipModuleGetFunction(&k->f_precalc, k->mod, "rfk_precalc"));
    RFK_HIP(hipModuleGetFunction(&k->f_batch, k->mod, k->opt.bin_queue ? "rfk_iterate_qbatch" : "rfk_iterate_batch"));
    *out = k.release();
    return RFK_OK;
}

int rfk_compile_check(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg,
                      size_t* code_bytes, char* log, size_t log_cap) {
    if (!db || !f) return fail(RFK_ERR_INVALID, "null argument");
    if (log && log_cap) log[0] = '\0';
    rfk::compile_options o;
    if (int rc = make_options(cfg, precision, o)) return rc;
    std::string src;
    try {
        auto L = rfk::make_layout(f->f, db->db);
        src = rfk::make_source(db->db, f->f, L, o);
    } catch (const std::exception& e) {
        return fail(RFK_ERR_PARSE, e.what());
    }
    auto cr = rfk::compile_hip(src, "rfk_flame", rfk::default_compile_flags());
    if (log && log_cap) {
        size_t n = std::min(log_cap - 1, cr.log.size());
        std::memcpy(log, cr.log.data(), n);
        log[n] = '\0';
    }
    if (!cr.ok) return fail(RFK_ERR_COMPILE, "hiprtc compile failed:\n" + cr.log);
    if (code_bytes) *code_bytes = cr.code.size();
    return RFK_OK;
}

int rfk_kernel_free(rfk_kernel* k) {
    if (k && k->mod) (void)hipModuleUnload(k->mod);
    delete k;
    return RFK_OK;
}

int rfk_kernel_size_reals(const rfk_kernel* k, size_t* out) {
    if (!k || !out) return fail(RFK_ERR_INVALID, "null argument");
    *out = size_t(k->L.size_reals);
    return RFK_OK;
}

int rfk_kernel_threads_per_block(const rfk_kernel* k, uint32_t* out) {
    if (!k || !out) return fail(RFK_ERR_INVALID, "null argument");
    *out = k->opt.tpb;
    return RFK_OK;
}

double rfk_kernel_compile_ms(const rfk_kernel* k) { return k ? k->compile_ms : 0.0; }

// ------------------------------------------------------------------ render

int rfk_fix_rotation(double* samples, size_t sample_size, size_t flame_size, const size_t* affine_indices,
                     size_t n_indices, int count) {
    if (!samples || (!affine_indices && n_indices) || count <= 0 || flame_size > sample_size)
        return fail(RFK_ERR_INVALID, "bad argument");
    rfk::fix_rotation(samples, sample_size, flame_size, affine_indices, n_indices, count);
    return RFK_OK;
}

// Everything of a warmup that does not allocate: clear the histogram and pass
// counters, fix rotations, Catmull-Rom + per-block sample evaluation,
// precalc, walker seeding and the fuse passes.
static int state_prepare(rfk_kernel* k, rfk_state* s, const double* samples, size_t n_reals, uint32_t seed,
                         uint32_t stream_id, uint32_t fuse) {
    const size_t sample_size = size_t(k->L.size_reals) + 768;
    if (n_reals % sample_size != 0 || n_reals / sample_size < 4)
        return fail(RFK_ERR_INVALID, "samples must hold >= 4 packed samples of " + std::to_string(sample_size) + " reals");
    const uint32_t nsamp = uint32_t(n_reals / sample_size), nseg = nsamp - 3;
    s->seed = seed;
    s->stream_id = stream_id;
    RFK_HIP(hipMemsetAsync(s->hist, 0, s->hist_words() * 4, s->stream));
    RFK_HIP(hipMemsetAsync(s->block_pass, 0, size_t(s->blocks) * 4, s->stream));
    if (s->wlast) RFK_HIP(hipMemsetAsync(s->wlast, 0xFF, size_t(s->walkers) * 4, s->stream));
    // sample scratch grows with the number of keyframe samples
    if (s->scratch_samples < nsamp) {
        if (s->d_samples) (void)hipFree(s->d_samples);
        if (s->d_segments) (void)hipFree(s->d_segments);
        s->d_samples = nullptr;
        s->d_segments = nullptr;
        s->scratch_samples = 0;
        RFK_HIP(hipMalloc(&s->d_samples, n_reals * sizeof(double)));
        RFK_HIP(hipMalloc(&s->d_segments, size_t(nseg) * sample_size * 4 * sizeof(double)));
        s->scratch_samples = nsamp;
    }
    // pinned staging copy of the samples; the previous upload from it must be done
    if (s->staging_reals < n_reals) {
        if (s->staging) {
            (void)hipEventSynchronize(s->ev_staging);
            (void)hipHostFree(s->staging);
        }
        s->staging = nullptr;
        s->staging_reals = 0;
        RFK_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->staging), n_reals * sizeof(double), hipHostMallocDefault));
        s->staging_reals = n_reals;
    } else {
        RFK_HIP(hipEventSynchronize(s->ev_staging));
    }
    std::memcpy(s->staging, samples, n_reals * sizeof(double));
    rfk::fix_rotation(s->staging, sample_size, size_t(k->L.size_reals), k->L.affine_indices.data(),
                      k->L.affine_indices.size(), int(nsamp));
    RFK_HIP(hipMemcpyAsync(s->d_samples, s->staging, n_reals * sizeof(double), hipMemcpyHostToDevice, s->stream));
    RFK_HIP(hipEventRecord(s->ev_staging, s->stream));
    RFK_HIP(rfk::sk::catmull(s->d_samples, uint32_t(sample_size), nseg, s->d_segments, s->stream));
    RFK_HIP(rfk::sk::eval_samples(s->d_segments, uint32_t(sample_size), s->size_reals, nseg, s->blocks, s->d_kind,
                                  s->precision, s->params, s->param_stride, s->palettes, s->stream));
    // per-sample precalc + jitter (generated), then walker seeding and fuse passes
    rfk_iter_args a = base_args(s);
    {
        void* argv[] = {&a};
        uint32_t tb = 64, g = (s->blocks + tb - 1) / tb;
        RFK_HIP(hipModuleLaunchKernel(k->f_precalc, g, 1, 1, tb, 1, 1, 0, s->stream, argv, nullptr));
    }
    if (int rc = reset_counters(s)) return rc;
    a.flags = RFK_FLAG_INIT;
    a.npasses = fuse;
    if (int rc = launch_iterate(s, a)) return rc;
    s->passes_done = fuse;
    s->quality = 0.0;
    return RFK_OK;
}

int rfk_warmup(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t width, uint32_t height,
               uint32_t seed, uint32_t fuse, uint32_t walkers, int temporal_multiplier, rfk_state** out) {
    return rfk_warmup_ex(k, stream, samples, n_reals, width, height, seed, 0u, fuse, walkers, temporal_multiplier, out);
}

int rfk_warmup_ex(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t width, uint32_t height,
                  uint32_t seed, uint32_t stream_id, uint32_t fuse, uint32_t walkers, int temporal_multiplier,
                  rfk_state** out) {
    if (!k || !samples || !out) return fail(RFK_ERR_INVALID, "null argument");
    if (width == 0 || height == 0 || width > 32768 || height > 32768) return fail(RFK_ERR_INVALID, "bad image size");
    if (temporal_multiplier < 1 || temporal_multiplier > 64) return fail(RFK_ERR_INVALID, "bad temporal multiplier");
    const size_t sample_size = size_t(k->L.size_reals) + 768;
    if (n_reals % sample_size != 0 || n_reals / sample_size < 4)
        return fail(RFK_ERR_INVALID, "samples must hold >= 4 packed samples of " + std::to_string(sample_size) + " reals");
    if (int rc = check_device()) return rc;

    auto s = std::make_unique<rfk_state>();
    s->k = k;
    s->width = width;
    s->height = height;
    s->tpb = k->opt.tpb;
    if (walkers == 0) walkers = 256u * 1024u;
    uint64_t base_blocks = (uint64_t(walkers) + s->tpb - 1) / s->tpb;
    uint64_t blocks = base_blocks * uint64_t(temporal_multiplier);
    if (blocks * s->tpb > (1ull << 31)) return fail(RFK_ERR_INVALID, "too many walkers");
    s->blocks = uint32_t(blocks);
    s->walkers = uint32_t(blocks * s->tpb);
    s->size_reals = uint32_t(k->L.size_reals);
    s->precision = k->opt.precision;
    const uint32_t align = s->precision ? 8u : 16u;  // 64-byte rows
    s->param_stride = (uint32_t(k->L.device_reals()) + align - 1) / align * align;
    s->seed = seed;
    s->stream_id = stream_id;
    if (stream) {
        s->stream = static_cast<hipStream_t>(stream);
    } else {
        RFK_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
    }
    int dev = 0;
    RFK_HIP(hipGetDevice(&dev));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
        s->wallclock_khz = double(khz);

    rfk_state* S = s.get();
    auto cleanup_on_error = [&](int rc) {
        free_state_buffers(S);
        S->own_stream = false;
        return rc;
    };
    const size_t elem = s->precision ? sizeof(double) : sizeof(float);
    const size_t npix = size_t(width) * height;
#define RFK_HIPC(expr)                                                                                    \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return cleanup_on_error(fail(RFK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    RFK_HIPC(hipMalloc(&s->params, size_t(s->blocks) * s->param_stride * elem));
    RFK_HIPC(dalloc(&s->palettes, size_t(s->blocks) * 256));
    RFK_HIPC(dalloc(&s->wx, s->walkers));
    RFK_HIPC(dalloc(&s->wy, s->walkers));
    RFK_HIPC(dalloc(&s->wc, s->walkers));
    if (k->L.chaos_base >= 0) RFK_HIPC(dalloc(&s->wlast, s->walkers));
    RFK_HIPC(dalloc(&s->block_pass, s->blocks));
    s->hist64 = k->opt.hist64;
    RFK_HIPC(dalloc(&s->hist, s->hist_words()));
    if (k->opt.bin_queue) {
        // chunks of whole passes of all walkers, at most about 2^27 records
        // (1 GiB) each; their record buffers come with the first bin call
        s->chunk_passes = std::max<uint32_t>(1u, uint32_t((1ull << 27) / s->walkers));
        const size_t q = size_t(s->chunk_passes) * s->walkers;
        const uint32_t ntiles = rfk::sk::queue_tiles(s->width, s->height);
        if (q >= (1ull << 32)) return cleanup_on_error(fail(RFK_ERR_INVALID, "too many walkers for queue binning"));
        RFK_HIPC(dalloc(&s->qcounts, size_t(ntiles) * s->blocks));
        RFK_HIPC(dalloc(&s->qtiles, size_t(ntiles) * 2));
        RFK_HIPC(dalloc(&s->qnseg, 4));
    }
    // the call's counters and its density statistics in one allocation, so
    // that one copy brings both to the host
    RFK_HIPC(dalloc(&s->counters, 8 + 98));
    s->dstats = s->counters + 8;
    RFK_HIPC(hipHostMalloc(reinterpret_cast<void**>(&s->counters_host), 128 * sizeof(unsigned long long),
                           hipHostMallocDefault));
    RFK_HIPC(hipHostMalloc(reinterpret_cast<void**>(&s->stop_host), 64, hipHostMallocMapped | hipHostMallocCoherent));
    *s->stop_host = 0;
    RFK_HIPC(hipHostGetDevicePointer(reinterpret_cast<void**>(&s->stop_dev), s->stop_host, 0));
    RFK_HIPC(hipEventCreate(&s->ev0));
    RFK_HIPC(hipEventCreate(&s->ev1));
    RFK_HIPC(hipEventCreateWithFlags(&s->ev_result, hipEventDisableTiming));
    RFK_HIPC(hipEventCreateWithFlags(&s->ev_staging, hipEventDisableTiming));
    std::vector<int8_t> kind(size_t(k->L.size_reals), 0);
    for (size_t ai : k->L.affine_indices)
        for (int c = 0; c < 4; c++) kind[ai + size_t(c)] = int8_t(1 + c);
    RFK_HIPC(hipMalloc(&s->d_kind, kind.size() + 16));
    RFK_HIPC(hipMemcpy(s->d_kind, kind.data(), kind.size(), hipMemcpyHostToDevice));
#undef RFK_HIPC
    if (int rc = state_prepare(k, S, samples, n_reals, seed, stream_id, fuse)) return cleanup_on_error(rc);
    *out = s.release();
    return RFK_OK;
}

int rfk_rewarm(rfk_kernel* k, rfk_state* s, const double* samples, size_t n_reals, uint32_t seed, uint32_t stream_id,
               uint32_t fuse) {
    if (!k || !s || !samples) return fail(RFK_ERR_INVALID, "null argument");
    if (s->k != k) return fail(RFK_ERR_INVALID, "state belongs to another kernel");
    if (s->pending) return fail(RFK_ERR_INVALID, "a bin call is pending on this state");
    return state_prepare(k, s, samples, n_reals, seed, stream_id, fuse);
}

int rfk_state_free(rfk_state* s) {
    if (!s) return RFK_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_state_buffers(s);
    delete s;
    return RFK_OK;
}

int rfk_bin_async(rfk_kernel* k, rfk_state* s, const rfk_bailout* bo) {
    if (!k || !s || !bo) return fail(RFK_ERR_INVALID, "null argument");
    if (s->k != k) return fail(RFK_ERR_INVALID, "state belongs to another kernel");
    if (s->pending) return fail(RFK_ERR_INVALID, "a bin call is already pending on this state");
    const bool has_time = bo->millis > 0.0, has_q = bo->quality > 0.0;
    if (bo->iters == 0 && !has_time && !has_q) return fail(RFK_ERR_INVALID, "no bailout set");
    rfk_iter_args a = base_args(s);
    uint64_t passes = bo->iters ? (bo->iters + s->walkers - 1) / s->walkers : 0;
    if (passes == 0) passes = 0x7FFFFFFFull;  // bounded by the time / quality bailouts
    if (passes > 0x7FFFFFFFull) return fail(RFK_ERR_INVALID, "iteration budget too large for one call");
    a.npasses = uint32_t(passes);
    a.flags = RFK_FLAG_DRAW;
    if (has_time || has_q) a.flags |= RFK_FLAG_BAILOUT;
    a.time_limit_ticks = has_time ? uint64_t(bo->millis * s->wallclock_khz) : 0;
    // atomic binning checks quality per block (its share), queue binning
    // against the whole call's hits
    const double q_total = bo->quality * double(s->width) * s->height * 255.0;
    a.quality_target = has_q ? uint64_t(k->opt.bin_queue ? q_total : q_total / s->blocks) : 0;
    *s->stop_host = 0;
    if (int rc = reset_counters(s)) return rc;
    RFK_HIP(hipEventRecord(s->ev0, s->stream));
    if (!k->opt.bin_queue) {
        if (int rc = launch_iterate(s, a)) return rc;
    } else if (int rc = ensure_queue(s, passes)) {
        return rc;
    } else if (int rc2 = bin_queued(s, a, passes, bo->iters != 0)) {
        return rc2;
    }
    RFK_HIP(hipEventRecord(s->ev1, s->stream));
    // density statistics after every bin, as the reference does
    // (flame_compiler.cpp:928-937), into the bins reset_counters cleared;
    // counters and statistics to the host in one copy
    RFK_HIP(rfk::sk::density_histogram(s->hist, s->hist64, size_t(s->width) * s->height, s->dstats, s->dstats + 97,
                                       s->stream));
    RFK_HIP(hipMemcpyAsync(s->counters_host, s->counters, (8 + 98) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s->stream));
    RFK_HIP(hipEventRecord(s->ev_result, s->stream));
    s->pending = true;
    s->pending_passes = passes;
    return RFK_OK;
}

int rfk_bin_wait(rfk_state* s, rfk_bin_result* out) {
    if (!s || !out) return fail(RFK_ERR_INVALID, "null argument");
    if (!s->pending) return fail(RFK_ERR_INVALID, "no bin call pending");
    // the call's own work only: what the caller enqueued behind it (density
    // estimation, tonemap) keeps running
    RFK_HIP(hipEventSynchronize(s->ev_result));
    s->pending = false;
    const unsigned long long* c = s->counters_host;
    const double bins = double(s->width) * s->height;
    std::memset(out, 0, sizeof *out);
    out->hit_sum = c[RFK_CNT_HITS];
    out->quality = double(c[RFK_CNT_HITS]) / (bins * 255.0);
    out->total_passes = c[RFK_CNT_PASSES];
    out->total_draws = c[RFK_CNT_DRAWS];
    out->bad_values = c[RFK_CNT_BAD];
    out->total_bins = uint64_t(bins);
    out->passes_per_thread = double(c[RFK_CNT_PASSES]) / s->walkers;
    out->elapsed_ms = c[RFK_CNT_STOP] > c[RFK_CNT_START] ? double(c[RFK_CNT_STOP] - c[RFK_CNT_START]) / s->wallclock_khz : 0.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s->ev0, s->ev1) == hipSuccess) out->kernel_ms = ms;
    out->max_density = double(c[8 + 97]);
    s->quality += out->quality;
    s->passes_done += uint64_t(out->passes_per_thread + 0.5);
    return RFK_OK;
}

int rfk_bin(rfk_kernel* k, rfk_state* s, const rfk_bailout* bo, rfk_bin_result* out) {
    if (int rc = rfk_bin_async(k, s, bo)) return rc;
    rfk_bin_result tmp;
    int rc = rfk_bin_wait(s, out ? out : &tmp);
    return rc;
}

int rfk_abort(rfk_state* s) {
    if (!s || !s->stop_host) return fail(RFK_ERR_INVALID, "null state");
    __atomic_store_n(s->stop_host, 1u, __ATOMIC_SEQ_CST);
    return RFK_OK;
}

int rfk_bin_points(rfk_state* s, const float* xyco, size_t n) {
    if (!s || (!xyco && n)) return fail(RFK_ERR_INVALID, "null argument");
    if (s->precision != 0) return fail(RFK_ERR_INVALID, "bin_points needs an f32 state");
    RFK_HIP(rfk::sk::bin_points(static_cast<const float*>(s->params), s->palettes, xyco, n, s->hist, s->hist64, s->width,
                                s->height, s->stream));
    RFK_HIP(hipStreamSynchronize(s->stream));
    return RFK_OK;
}

int rfk_trace(rfk_kernel* k, rfk_state* s, uint32_t passes, float* trace, uint32_t* xform_ids) {
    if (!k || !s || !trace) return fail(RFK_ERR_INVALID, "null argument");
    if (s->k != k) return fail(RFK_ERR_INVALID, "state belongs to another kernel");
    rfk_iter_args a = base_args(s);
    a.npasses = passes;
    a.flags = RFK_FLAG_TRACE;
    a.trace = trace;
    a.trace_xid = xform_ids;
    if (int rc = reset_counters(s)) return rc;
    if (int rc = launch_iterate(s, a)) return rc;
    RFK_HIP(hipStreamSynchronize(s->stream));
    s->passes_done += passes;
    return RFK_OK;
}

int rfk_state_get_info(const rfk_state* s, rfk_state_info* out) {
    if (!s || !out) return fail(RFK_ERR_INVALID, "null argument");
    out->width = s->width;
    out->height = s->height;
    out->walkers = s->walkers;
    out->threads_per_block = s->tpb;
    out->blocks = s->blocks;
    out->size_reals = s->size_reals;
    out->param_stride = s->param_stride;
    out->seed = s->seed;
    out->passes_done = s->passes_done;
    out->histogram = s->hist;
    out->params = static_cast<float*>(s->params);
    out->palettes = s->palettes;
    out->walker_x = s->wx;
    out->walker_y = s->wy;
    out->walker_c = s->wc;
    out->stream = s->stream;
    out->quality = s->quality;
    out->hist_bits = s->hist64 ? 64u : 32u;
    out->walker_last = s->wlast;
    return RFK_OK;
}

int rfk_state_clear_histogram(rfk_state* s) {
    if (!s) return fail(RFK_ERR_INVALID, "null state");
    RFK_HIP(hipMemsetAsync(s->hist, 0, s->hist_words() * 4, s->stream));
    RFK_HIP(hipStreamSynchronize(s->stream));
    s->quality = 0.0;
    return RFK_OK;
}

// ---------------------------------------------------------------- autotune

int rfk_autotune(const rfk_db* db, const rfk_flame* f, int precision, const double* samples, size_t n_reals,
                 uint32_t width, uint32_t height, uint64_t probe_iters, rfk_tune_candidate* cands, size_t n_cands,
                 size_t* best) {
    if (!db || !f || !samples || !cands || !n_cands || !best) return fail(RFK_ERR_INVALID, "null argument");
    if (probe_iters == 0) return fail(RFK_ERR_INVALID, "probe_iters must be > 0");
    double best_rate = -1.0;
    size_t best_i = n_cands;
    for (size_t i = 0; i < n_cands; i++) {
        rfk_tune_candidate& c = cands[i];
        c.iters_per_s = 0.0;
        c.compile_ms = 0.0;
        rfk_exec_config cfg{c.threads_per_block, c.atomic_width, c.bin_tile_log2, 0u, c.bin_queue};
        rfk_kernel* k = nullptr;
        c.status = rfk_compile(db, f, precision, &cfg, &k, nullptr, 0);
        if (c.status != RFK_OK) continue;
        c.compile_ms = k->compile_ms;
        rfk_state* s = nullptr;
        c.status = rfk_warmup(k, nullptr, samples, n_reals, width, height, 1u, 20u, c.walkers, 1, &s);
        if (c.status == RFK_OK) {
            rfk_bailout bo{std::max<uint64_t>(probe_iters / 4, 1), 0.0, 0.0};
            rfk_bin_result r{};
            c.status = rfk_bin(k, s, &bo, &r);  // untimed: clocks and caches settle
            if (c.status == RFK_OK) {
                bo.iters = probe_iters;
                c.status = rfk_bin(k, s, &bo, &r);
            }
            if (c.status == RFK_OK && r.kernel_ms > 0.0) c.iters_per_s = double(r.total_passes) / (r.kernel_ms * 1e-3);
            rfk_state_free(s);
        }
        rfk_kernel_free(k);
        if (c.status == RFK_OK && c.iters_per_s > best_rate) {
            best_rate = c.iters_per_s;
            best_i = i;
        }
    }
    if (best_i == n_cands) return fail(RFK_ERR_INVALID, "no candidate configuration ran");
    *best = best_i;
    return RFK_OK;
}

// ------------------------------------------------------------------- batch

static void free_batch(rfk_batch* b) {
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(b->params); f(b->palettes); f(b->hist); f(b->d_item_frame); f(b->d_item_block); f(b->d_frame_row0);
    f(b->d_frame_blocks); f(b->d_frame_passes); f(b->d_queue); f(b->d_frame_counters); f(b->d_row_stream);
    f(b->d_row_local); f(b->d_slots); f(b->d_slot_frame); f(b->qcounts); f(b->qtiles); f(b->qnseg); f(b->wx);
    f(b->wy); f(b->wc); f(b->records); f(b->sorted); f(b->qsegments);
    if (b->h_frame_counters) (void)hipHostFree(b->h_frame_counters);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
}

// Queue-binned batch: frame f's binned passes are cut into slices of
// slice_passes; a chunk launch runs up to slots_per_chunk slices of different
// frames (every walker block of each) and then sorts and sums the chunk's
// records into the frames' histograms. Chunks take the frames with the most
// slices left, so the longest frames set the chunk count only when they
// must. Sizes: enough walker blocks per launch to fill the device (2048
// workgroups) and about 2^27 records (1 GiB) per chunk.
static int plan_queue_batch(rfk_batch* B) {
    const uint32_t bpf = B->blocks_per_frame, frames = B->frames;
    const uint64_t wf = uint64_t(bpf) * B->tpb;
    const uint32_t S = std::min<uint32_t>(frames, std::max<uint32_t>(1u, (2048u + bpf - 1) / bpf));
    const uint64_t P = std::max<uint64_t>(1, (1ull << 27) / (uint64_t(S) * wf));
    const uint64_t records = uint64_t(S) * wf * P;
    if (records >= (1ull << 32)) return fail(RFK_ERR_INVALID, "too many walkers per frame for queue binning");
    B->queue = true;
    B->slots_per_chunk = S;
    B->slice_passes = uint32_t(P);
    B->tiles_per_frame = rfk::sk::queue_tiles(B->width, B->height);
    std::vector<uint32_t> left(frames), next(frames, 0), slots, slot_frame, pick;
    uint64_t remaining = 0;
    for (uint32_t f = 0; f < frames; f++) {
        left[f] = uint32_t((B->frame_passes[f] + P - 1) / P);
        remaining += left[f];
    }
    while (remaining) {
        pick.clear();
        for (uint32_t f = 0; f < frames; f++)
            if (left[f]) pick.push_back(f);
        std::stable_sort(pick.begin(), pick.end(), [&](uint32_t x, uint32_t y) { return left[x] > left[y]; });
        if (pick.size() > S) pick.resize(S);
        rfk_batch::chunk c{uint32_t(slot_frame.size()), uint32_t(pick.size()), 0u};
        for (uint32_t f : pick) {
            const uint32_t sl = next[f]++;
            const uint64_t first = uint64_t(sl) * P;
            const uint32_t n = uint32_t(std::min<uint64_t>(P, B->frame_passes[f] - first));
            // passes are numbered as in one render: fuse passes, then the binned ones
            slots.insert(slots.end(), {f, sl ? uint32_t(B->fuse + first) : 0u, n, sl ? 0u : B->fuse});
            slot_frame.push_back(f);
            c.passes = std::max(c.passes, n);
            left[f]--;
            remaining--;
        }
        B->chunks.push_back(c);
    }
    const uint32_t ctiles = S * B->tiles_per_frame;
    B->max_segments = rfk::sk::queue_max_segments(records, ctiles);
    B->nitems = uint32_t(slot_frame.size()) * bpf;
    B->grid = S * bpf;
    RFK_HIP(dalloc(&B->d_slots, slots.size()));
    RFK_HIP(dalloc(&B->d_slot_frame, slot_frame.size()));
    RFK_HIP(dalloc(&B->wx, size_t(frames) * wf));
    RFK_HIP(dalloc(&B->wy, size_t(frames) * wf));
    RFK_HIP(dalloc(&B->wc, size_t(frames) * wf));
    RFK_HIP(hipMalloc(&B->records, records * 8));
    RFK_HIP(hipMalloc(&B->sorted, records * 8));
    RFK_HIP(dalloc(&B->qcounts, size_t(ctiles) * bpf));
    RFK_HIP(dalloc(&B->qtiles, size_t(ctiles) * 2));
    RFK_HIP(hipMalloc(&B->qsegments, size_t(B->max_segments) * 16));
    RFK_HIP(dalloc(&B->qnseg, 1));
    RFK_HIP(hipMemcpyAsync(B->d_slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, B->stream));
    RFK_HIP(hipMemcpyAsync(B->d_slot_frame, slot_frame.data(), slot_frame.size() * 4, hipMemcpyHostToDevice,
                           B->stream));
    RFK_HIP(hipStreamSynchronize(B->stream));
    return RFK_OK;
}

int rfk_batch_create(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t frames,
                     const uint64_t* iters, uint32_t width, uint32_t height, uint32_t walkers_per_frame, uint32_t seed,
                     uint32_t fuse, rfk_batch** out) {
    if (!k || !samples || !iters || !out || frames == 0) return fail(RFK_ERR_INVALID, "bad argument");
    if (k->L.chaos_base >= 0) return fail(RFK_ERR_INVALID, "batches need a kernel without xaos");
    if (width == 0 || height == 0 || width > 32768 || height > 32768) return fail(RFK_ERR_INVALID, "bad image size");
    const size_t sample_size = size_t(k->L.size_reals) + 768;
    if (n_reals % sample_size != 0 || n_reals / sample_size < 4)
        return fail(RFK_ERR_INVALID, "samples must hold >= 4 packed keyframe samples");
    if (int rc = check_device()) return rc;
    auto b = std::make_unique<rfk_batch>();
    rfk_batch* B = b.get();
    B->k = k;
    B->frames = frames;
    B->width = width;
    B->height = height;
    B->tpb = k->opt.tpb;
    B->blocks_per_frame = std::max<uint32_t>(1, (walkers_per_frame + B->tpb - 1) / B->tpb);
    const uint64_t total = uint64_t(B->blocks_per_frame) * frames;
    if (total > (1u << 24)) return fail(RFK_ERR_INVALID, "too many blocks in the batch");
    B->total_blocks = uint32_t(total);
    const uint32_t align = k->opt.precision ? 8u : 16u;
    B->param_stride = (uint32_t(k->L.device_reals()) + align - 1) / align * align;
    B->seed = seed;
    B->fuse = fuse;
    const uint64_t walkers_f = uint64_t(B->blocks_per_frame) * B->tpb;
    B->frame_passes.resize(frames);
    for (uint32_t f = 0; f < frames; f++) {
        uint64_t p = (iters[f] + walkers_f - 1) / walkers_f;
        if (p == 0 || p > 0x7FFFFFF0ull) return fail(RFK_ERR_INVALID, "bad frame iteration budget");
        B->frame_passes[f] = uint32_t(p);
    }
    auto cleanup = [&](int rc) {
        free_batch(B);
        B->own_stream = false;
        return rc;
    };
#define RFK_HIPB(expr)                                                                                    \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return cleanup(fail(RFK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    if (stream) {
        B->stream = static_cast<hipStream_t>(stream);
    } else {
        RFK_HIPB(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        B->own_stream = true;
    }
    const size_t elem = k->opt.precision ? sizeof(double) : sizeof(float);
    const size_t fpix = size_t(width) * height;
    RFK_HIPB(hipMalloc(&B->params, size_t(B->total_blocks) * B->param_stride * elem));
    RFK_HIPB(dalloc(&B->palettes, size_t(B->total_blocks) * 256));
    B->hist64 = k->opt.hist64;
    RFK_HIPB(dalloc(&B->hist, B->frame_words() * frames));
    RFK_HIPB(dalloc(&B->d_frame_row0, frames));
    RFK_HIPB(dalloc(&B->d_frame_blocks, frames));
    RFK_HIPB(dalloc(&B->d_frame_passes, frames));
    RFK_HIPB(dalloc(&B->d_frame_counters, size_t(frames) * 4));
    RFK_HIPB(dalloc(&B->d_queue, 4));
    RFK_HIPB(dalloc(&B->d_item_frame, B->total_blocks));
    RFK_HIPB(dalloc(&B->d_item_block, B->total_blocks));
    RFK_HIPB(dalloc(&B->d_row_stream, B->total_blocks));
    RFK_HIPB(dalloc(&B->d_row_local, B->total_blocks));
    RFK_HIPB(hipHostMalloc(reinterpret_cast<void**>(&B->h_frame_counters), size_t(frames) * 4 * 8, hipHostMallocDefault));
    RFK_HIPB(hipEventCreate(&B->ev0));
    RFK_HIPB(hipEventCreate(&B->ev1));

    // rows: frame f owns rows [f * bpf, (f+1) * bpf); block b of frame f sits at
    // global time (f + b / bpf) * nseg / frames (one frame interval of motion blur)
    const uint32_t nsamp = uint32_t(n_reals / sample_size), nseg = nsamp - 3, bpf = B->blocks_per_frame;
    std::vector<uint32_t> row0(frames), fblocks(frames, bpf), row_stream(B->total_blocks), row_local(B->total_blocks);
    std::vector<double> times(B->total_blocks);
    for (uint32_t f = 0; f < frames; f++) {
        row0[f] = f * bpf;
        for (uint32_t bl = 0; bl < bpf; bl++) {
            times[f * bpf + bl] = (double(f) + double(bl) / bpf) * nseg / frames;
            row_stream[f * bpf + bl] = f;
            row_local[f * bpf + bl] = bl;
        }
    }
    // work items, longest frames first so the persistent grid drains evenly
    std::vector<uint32_t> order(frames);
    for (uint32_t f = 0; f < frames; f++) order[f] = f;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return B->frame_passes[x] > B->frame_passes[y]; });
    std::vector<uint32_t> item_frame, item_block;
    for (uint32_t f : order)
        for (uint32_t bl = 0; bl < bpf; bl++) {
            item_frame.push_back(f);
            item_block.push_back(bl);
        }
    B->nitems = uint32_t(item_frame.size());
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, B->stream);
    };
    RFK_HIPB(up(B->d_frame_row0, row0.data(), frames * 4));
    RFK_HIPB(up(B->d_frame_blocks, fblocks.data(), frames * 4));
    RFK_HIPB(up(B->d_frame_passes, B->frame_passes.data(), frames * 4));
    RFK_HIPB(up(B->d_item_frame, item_frame.data(), item_frame.size() * 4));
    RFK_HIPB(up(B->d_item_block, item_block.data(), item_block.size() * 4));
    RFK_HIPB(up(B->d_row_stream, row_stream.data(), row_stream.size() * 4));
    RFK_HIPB(up(B->d_row_local, row_local.data(), row_local.size() * 4));

    // keyframes -> segments -> per-row parameters and palettes -> precalc
    std::vector<double> host(samples, samples + n_reals);
    rfk::fix_rotation(host.data(), sample_size, size_t(k->L.size_reals), k->L.affine_indices.data(),
                      k->L.affine_indices.size(), int(nsamp));
    std::vector<int8_t> kind(size_t(k->L.size_reals), 0);
    for (size_t ai : k->L.affine_indices)
        for (int c = 0; c < 4; c++) kind[ai + size_t(c)] = int8_t(1 + c);
    double *d_samples = nullptr, *d_segments = nullptr, *d_times = nullptr;
    int8_t* d_kind = nullptr;
    RFK_HIPB(hipMalloc(&d_samples, n_reals * sizeof(double)));
    RFK_HIPB(hipMalloc(&d_segments, size_t(nseg) * sample_size * 4 * sizeof(double)));
    RFK_HIPB(hipMalloc(&d_times, times.size() * sizeof(double)));
    RFK_HIPB(hipMalloc(&d_kind, kind.size() + 16));
    RFK_HIPB(up(d_samples, host.data(), n_reals * sizeof(double)));
    RFK_HIPB(up(d_times, times.data(), times.size() * sizeof(double)));
    RFK_HIPB(up(d_kind, kind.data(), kind.size()));
    RFK_HIPB(rfk::sk::catmull(d_samples, uint32_t(sample_size), nseg, d_segments, B->stream));
    RFK_HIPB(rfk::sk::eval_samples(d_segments, uint32_t(sample_size), uint32_t(k->L.size_reals), nseg, B->total_blocks,
                                   d_kind, k->opt.precision, B->params, B->param_stride, B->palettes, B->stream,
                                   d_times));
    rfk_iter_args a;
    std::memset(&a, 0, sizeof a);
    a.params = static_cast<const float*>(B->params);
    a.param_stride = B->param_stride;
    a.seed = seed;
    a.nsamples = B->total_blocks;
    a.row_stream = B->d_row_stream;
    a.row_local = B->d_row_local;
    {
        void* argv[] = {&a};
        uint32_t tb = 64, g = (B->total_blocks + tb - 1) / tb;
        RFK_HIPB(hipModuleLaunchKernel(k->f_precalc, g, 1, 1, tb, 1, 1, 0, B->stream, argv, nullptr));
    }
    RFK_HIPB(hipStreamSynchronize(B->stream));
    (void)hipFree(d_samples);
    (void)hipFree(d_segments);
    (void)hipFree(d_times);
    (void)hipFree(d_kind);
    // persistent grid: enough groups to fill every CU at the kernel's occupancy
    int dev = 0, cus = 0, per_cu = 0;
    RFK_HIPB(hipGetDevice(&dev));
    RFK_HIPB(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k->f_batch, int(B->tpb), 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    B->grid = std::min<uint32_t>(B->nitems, uint32_t(cus) * uint32_t(per_cu));
    if (k->opt.bin_queue) {
        if (int rc = plan_queue_batch(B)) return cleanup(rc);
    }
#undef RFK_HIPB
    *out = b.release();
    return RFK_OK;
}

int rfk_batch_free(rfk_batch* b) {
    if (!b) return RFK_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
    return RFK_OK;
}

static int batch_enqueue(rfk_batch* B) {
    rfk_kernel* k = B->k;
    const size_t fpix = size_t(B->width) * B->height;
    // clears are kernels, not memsets, so the sequence can live in a graph
    RFK_HIP(rfk::sk::fill_u32(B->hist, 0u, B->frame_words() * B->frames, B->stream));
    RFK_HIP(rfk::sk::fill_u32(B->d_queue, 0u, 1, B->stream));
    RFK_HIP(rfk::sk::fill_u32(reinterpret_cast<uint32_t*>(B->d_frame_counters), 0u, size_t(B->frames) * 8, B->stream));
    rfk_iter_args a;
    std::memset(&a, 0, sizeof a);
    a.params = static_cast<const float*>(B->params);
    a.palettes = B->palettes;
    a.hist = B->hist;
    a.width = B->width;
    a.height = B->height;
    a.param_stride = B->param_stride;
    a.seed = B->seed;
    a.check_every = 16;
    rfk_batch_args ba;
    std::memset(&ba, 0, sizeof ba);
    ba.queue = B->d_queue;
    ba.item_frame = B->d_item_frame;
    ba.item_block = B->d_item_block;
    ba.frame_row0 = B->d_frame_row0;
    ba.frame_blocks = B->d_frame_blocks;
    ba.frame_passes = B->d_frame_passes;
    ba.frame_counters = reinterpret_cast<u64*>(B->d_frame_counters);
    ba.nitems = B->nitems;
    ba.fuse = B->fuse;
    ba.frame_pixels = uint32_t(fpix);
    if (B->queue) {
        a.wx = B->wx;
        a.wy = B->wy;
        a.wc = B->wc;
        a.queue = static_cast<u64*>(B->records);
        a.qcounts = B->qcounts;
        a.tiles_x = rfk::sk::queue_tiles_x(B->width);
        a.ntiles = B->tiles_per_frame;
        rfk_qbatch_args q;
        std::memset(&q, 0, sizeof q);
        q.slots = B->d_slots;
        q.frame_counters = reinterpret_cast<u64*>(B->d_frame_counters);
        q.cols = B->blocks_per_frame;
        q.frame_pixels = uint32_t(fpix);
        void* qargv[] = {&a, &q};
        for (const auto& c : B->chunks) {
            q.slot0 = c.first;
            q.passes = c.passes;
            const uint32_t parts = c.nslots * B->blocks_per_frame;
            RFK_HIP(hipModuleLaunchKernel(k->f_batch, parts, 1, 1, B->tpb, 1, 1, a.ntiles * 4u, B->stream, qargv,
                                          nullptr));
            RFK_HIP(rfk::sk::queue_bin_groups(static_cast<const unsigned long long*>(B->records),
                                              size_t(c.passes) * B->tpb, parts, B->blocks_per_frame,
                                              static_cast<unsigned long long*>(B->sorted), B->qcounts, B->qtiles,
                                              B->qsegments, B->max_segments, B->qnseg, B->width, B->height,
                                              B->d_slot_frame + c.first, fpix * 4, B->hist, B->hist64,
                                              k->opt.atomic_width == 64, B->stream));
        }
        return RFK_OK;
    }
    void* argv[] = {&a, &ba};
    RFK_HIP(hipModuleLaunchKernel(k->f_batch, B->grid, 1, 1, B->tpb, 1, 1, 0, B->stream, argv, nullptr));
    return RFK_OK;
}

int rfk_batch_run(rfk_batch* B, int use_graph, double* kernel_ms) {
    if (!B) return fail(RFK_ERR_INVALID, "null batch");
    if (use_graph && !B->exec) {
        RFK_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue(B);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(B->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(RFK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        B->graph = g;
        RFK_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    }
    RFK_HIP(hipEventRecord(B->ev0, B->stream));
    if (use_graph) {
        RFK_HIP(hipGraphLaunch(B->exec, B->stream));
    } else if (int rc = batch_enqueue(B)) {
        return rc;
    }
    RFK_HIP(hipEventRecord(B->ev1, B->stream));
    RFK_HIP(hipMemcpyAsync(B->h_frame_counters, B->d_frame_counters, size_t(B->frames) * 32, hipMemcpyDeviceToHost,
                           B->stream));
    RFK_HIP(hipStreamSynchronize(B->stream));
    RFK_HIP(hipEventElapsedTime(&B->last_ms, B->ev0, B->ev1));
    if (kernel_ms) *kernel_ms = B->last_ms;
    return RFK_OK;
}

int rfk_batch_frame(const rfk_batch* B, uint32_t frame, rfk_batch_frame_info* out) {
    if (!B || !out || frame >= B->frames) return fail(RFK_ERR_INVALID, "bad argument");
    const unsigned long long* c = B->h_frame_counters + size_t(frame) * 4;
    const double bins = double(B->width) * B->height;
    out->histogram = B->hist + B->frame_words() * frame;
    out->quality = double(c[0]) / (bins * 255.0);
    out->total_draws = c[1];
    out->total_passes = c[2];
    out->bad_values = c[3];
    out->passes_per_walker = B->frame_passes[frame];
    out->walkers = B->blocks_per_frame * B->tpb;
    out->hist_bits = B->hist64 ? 64u : 32u;
    const size_t elem = B->k->opt.precision ? sizeof(double) : sizeof(float);
    const size_t row0 = size_t(frame) * B->blocks_per_frame;
    out->params = static_cast<const char*>(B->params) + row0 * B->param_stride * elem;
    out->palettes = B->palettes + row0 * 256;
    out->param_stride = B->param_stride;
    out->blocks = B->blocks_per_frame;
    out->seed = B->seed;
    out->stream_id = frame;
    out->fuse = B->fuse;
    return RFK_OK;
}

int rfk_batch_info(const rfk_batch* B, uint32_t* frames, uint32_t* width, uint32_t* height, uint32_t* grid,
                   uint32_t* items) {
    if (!B) return fail(RFK_ERR_INVALID, "null batch");
    if (frames) *frames = B->frames;
    if (width) *width = B->width;
    if (height) *height = B->height;
    if (grid) *grid = B->grid;
    if (items) *items = B->nitems;
    return RFK_OK;
}

// ------------------------------------------------------------ image kernels

int rfk_density_histogram(rfk_state* s, uint64_t* bins97, double* max_density) {
    if (!s) return fail(RFK_ERR_INVALID, "null state");
    RFK_HIP(hipMemsetAsync(s->dstats, 0, 98 * sizeof(unsigned long long), s->stream));
    RFK_HIP(rfk::sk::density_histogram(s->hist, s->hist64, size_t(s->width) * s->height, s->dstats, s->dstats + 97,
                                       s->stream));
    RFK_HIP(hipMemcpyAsync(s->counters_host + 8, s->dstats, 98 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           s->stream));
    RFK_HIP(hipStreamSynchronize(s->stream));
    if (bins97) std::memcpy(bins97, s->counters_host + 8, 97 * sizeof(uint64_t));
    if (max_density) *max_density = double(s->counters_host[8 + 97]);
    return RFK_OK;
}

int rfk_density_estimate(rfk_state* s, float max_radius, float min_radius, float curve, float* out_rgba) {
    if (!s || !out_rgba) return fail(RFK_ERR_INVALID, "null argument");
    if (!(max_radius >= 0.0f) || max_radius > 20.0f || !(min_radius >= 0.0f) || min_radius > 20.0f || !(curve >= 0.0f))
        return fail(RFK_ERR_INVALID, "radii must lie in [0, 20] and curve >= 0");
    const float hmax = std::max(max_radius, min_radius);
    const int nlevels = int(hmax * 16.0f + 0.5f) + 1;
    const int radius = (nlevels - 1 + 15) / 16;
    constexpr int kMaxLevels = 20 * 16 + 1;
    const size_t npix = size_t(s->width) * s->height;
    if (!s->de_levels) RFK_HIP(dalloc(&s->de_levels, npix));
    if (!s->de_inv_z) RFK_HIP(dalloc(&s->de_inv_z, size_t(kMaxLevels)));
    if (max_radius != s->de_max_radius || min_radius != s->de_min_radius) {
        // per-level normalisation 1/Z(h), Z(h) = sum over |v| <= h of exp(-2|v|^2/h^2)
        std::vector<float> inv_z(static_cast<size_t>(nlevels), 0.0f);
        for (int l = 0; l < nlevels; l++) {
            double h = l / 16.0, z = 0.0;
            for (int dy = -radius; dy <= radius; dy++)
                for (int dx = -radius; dx <= radius; dx++) {
                    double d2 = double(dx * dx + dy * dy);
                    if (d2 > h * h) continue;
                    z += d2 == 0.0 ? 1.0 : std::exp(-2.0 * d2 / (h * h));
                }
            inv_z[size_t(l)] = float(1.0 / z);
        }
        // the table may still be read by an earlier estimate on this stream
        RFK_HIP(hipStreamSynchronize(s->stream));
        RFK_HIP(hipMemcpy(s->de_inv_z, inv_z.data(), inv_z.size() * sizeof(float), hipMemcpyHostToDevice));
        s->de_max_radius = max_radius;
        s->de_min_radius = min_radius;
    }
    // enqueued on the state's stream; the call returns before it runs
    RFK_HIP(rfk::sk::density_estimate(s->hist, s->hist64, s->width, s->height, max_radius, min_radius, curve,
                                      s->de_inv_z, nlevels, radius, s->de_levels, out_rgba, s->stream));
    return RFK_OK;
}

static int tonemap_common(const void* in, int in_kind, size_t npix, void* out, int fmt, const rfk_tonemap_args* a,
                          double state_quality, hipStream_t stream) {
    if (!in || !out || !a) return fail(RFK_ERR_INVALID, "null argument");
    if (fmt < 0 || fmt > 4) return fail(RFK_ERR_INVALID, "bad pixel format");
    double q = a->quality > 0 ? a->quality : state_quality;
    if (!(q > 0)) return fail(RFK_ERR_INVALID, "quality must be positive (render first)");
    // scale constant as tonemapper.cpp:52: 10^(-log10(q) - 0.5) in float
    float k = std::pow(10.0f, -std::log10(float(q)) - 0.5f);
    RFK_HIP(rfk::sk::tonemap(in, in_kind, npix, out, fmt, k, float(a->brightness), float(a->gamma), float(a->vibrancy),
                             stream));
    return RFK_OK;
}

int rfk_tonemap(rfk_state* s, void* out, int pixel_format, const rfk_tonemap_args* a) {
    if (!s) return fail(RFK_ERR_INVALID, "null state");
    return tonemap_common(s->hist, s->hist64 ? 2 : 1, size_t(s->width) * s->height, out, pixel_format, a, s->quality,
                          s->stream);
}

int rfk_tonemap_features(rfk_state* s, const rfk_tonemap_args* a, void* feat, float* base) {
    if (!s || !a || !feat || !base) return fail(RFK_ERR_INVALID, "null argument");
    const size_t npix = size_t(s->width) * s->height;
    if (a->quality > 0) {
        const float k = std::pow(10.0f, -std::log10(float(a->quality)) - 0.5f);  // as tonemap_common
        RFK_HIP(rfk::sk::tonemap_features(s->hist, s->hist64, npix, k, nullptr, float(a->brightness),
                                          float(a->gamma), float(a->vibrancy), feat, base, s->stream));
    } else {
        // the quality of the last bin call enqueued on the state, read on
        // the device: no wait for the call
        RFK_HIP(rfk::sk::tonemap_features(s->hist, s->hist64, npix, 0.0f, s->counters + RFK_CNT_HITS,
                                          float(a->brightness), float(a->gamma), float(a->vibrancy), feat, base,
                                          s->stream));
    }
    return RFK_OK;
}

int rfk_tonemap_image(const float* in_rgba, uint32_t width, uint32_t height, void* out, int pixel_format,
                      const rfk_tonemap_args* a, void* stream) {
    return tonemap_common(in_rgba, 0, size_t(width) * height, out, pixel_format, a, 0.0,
                          static_cast<hipStream_t>(stream));
}

int rfk_convert_image(const void* in, int in_format, void* out, int out_format, uint32_t width, uint32_t height,
                      int flags, void* stream) {
    if (!in || !out) return fail(RFK_ERR_INVALID, "null argument");
    if (in_format != RFK_PIXEL_HALF3 && in_format != RFK_PIXEL_HALF4)
        return fail(RFK_ERR_INVALID, "conversion reads half3 or half4 images");
    if (out_format < RFK_PIXEL_FLOAT4 || out_format > RFK_PIXEL_RGB8) return fail(RFK_ERR_INVALID, "bad pixel format");
    if (flags & ~(RFK_CONVERT_FLIP | RFK_CONVERT_SWAP)) return fail(RFK_ERR_INVALID, "unknown conversion flags");
    RFK_HIP(rfk::sk::convert_image(in, in_format == RFK_PIXEL_HALF4 ? 4 : 3, out, out_format, width, height,
                                   (flags & RFK_CONVERT_FLIP) != 0, (flags & RFK_CONVERT_SWAP) != 0,
                                   static_cast<hipStream_t>(stream)));
    return RFK_OK;
}

int rfk_density_hdr(const void* tonemapped, int pixel_format, size_t npix, const uint64_t* bins97, float* out_rgba,
                    void* stream) {
    if (!tonemapped || !bins97 || !out_rgba) return fail(RFK_ERR_INVALID, "null argument");
    if (pixel_format != RFK_PIXEL_FLOAT4 && pixel_format != RFK_PIXEL_HALF4)
        return fail(RFK_ERR_INVALID, "density_hdr reads float4 or half4 images");
    hipStream_t st = static_cast<hipStream_t>(stream);
    unsigned long long* d_bins = nullptr;
    RFK_HIP(hipMalloc(&d_bins, 97 * sizeof(unsigned long long)));
    hipError_t e = hipMemcpyAsync(d_bins, bins97, 97 * sizeof(unsigned long long), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = rfk::sk::density_hdr(tonemapped, pixel_format == RFK_PIXEL_HALF4, npix, d_bins, out_rgba, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_bins);
    if (e != hipSuccess) return fail(RFK_ERR_HIP, std::string("density_hdr: ") + hipGetErrorString(e));
    return RFK_OK;
}

int rfk_catmull_coefficients(const double* samples, uint32_t sample_size, uint32_t nseg, double* segments,
                             void* stream) {
    if (!samples || !segments) return fail(RFK_ERR_INVALID, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    RFK_HIP(rfk::sk::catmull(samples, sample_size, nseg, segments, st));
    RFK_HIP(hipStreamSynchronize(st));
    return RFK_OK;
}

int rfk_atomic_peak(uint64_t footprint_bytes, uint64_t increments, double* increments_per_s) {
    if (!increments_per_s || footprint_bytes < 16) return fail(RFK_ERR_INVALID, "bad argument");
    if (int rc = check_device()) return rc;
    const uint64_t nbins = footprint_bytes / 16;
    uint32_t* buf = nullptr;
    RFK_HIP(hipMalloc(&buf, nbins * 16));
    const uint32_t block = 256, grid = 256 * 32;
    const uint64_t per_launch_threads = uint64_t(block) * grid;
    uint32_t iters = uint32_t(std::max<uint64_t>(1, increments / 4 / per_launch_threads));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipError_t e = hipMemset(buf, 0, nbins * 16);
    if (e == hipSuccess) e = rfk::sk::atomic_peak(buf, nbins, grid, block, 1, 1u, nullptr);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = rfk::sk::atomic_peak(buf, nbins, grid, block, iters, 7u, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(RFK_ERR_HIP, hipGetErrorString(e));
    *increments_per_s = double(per_launch_threads) * iters * 4.0 / (ms * 1e-3);
    return RFK_OK;
}

// ------------------------------------------------------------ device helpers

int rfk_device_count(int* n) {
    if (!n) return fail(RFK_ERR_INVALID, "null argument");
    *n = 0;
    if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
    return RFK_OK;
}
int rfk_malloc(void** p, size_t bytes) {
    RFK_HIP(hipMalloc(p, bytes));
    return RFK_OK;
}
int rfk_free(void* p) {
    RFK_HIP(hipFree(p));
    return RFK_OK;
}
// The helpers wait for all device work first: states run on their own
// non-blocking streams, which the null stream does not order against.
int rfk_memcpy_htod(void* dst, const void* src, size_t bytes) {
    RFK_HIP(hipDeviceSynchronize());
    RFK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RFK_OK;
}
int rfk_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    RFK_HIP(hipDeviceSynchronize());
    RFK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RFK_OK;
}
int rfk_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    RFK_HIP(hipDeviceSynchronize());
    RFK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return RFK_OK;
}
int rfk_memset(void* dst, int value, size_t bytes) {
    RFK_HIP(hipDeviceSynchronize());
    RFK_HIP(hipMemset(dst, value, bytes));
    return RFK_OK;
}
int rfk_synchronize(void) {
    RFK_HIP(hipDeviceSynchronize());
    return RFK_OK;
}

}  // extern "C"
