/*
 * refraktor_amd — C ABI of the MI355X-native fractal-flame renderer.
 *
 * Every entry point returns an int status (RFK_OK = 0, negative on error) and
 * never exits the process; the last error text of the calling thread is
 * available from rfk_last_error(). Handles are opaque; memory behind them is
 * owned by the library. Device pointers are plain `void*`/typed pointers into
 * HBM, host pointers are marked "host".
 *
 * Each declaration cites the librefrakt interface it replaces
 * (paths relative to the untrioctium/refraktor tree).
 */
#ifndef REFRAKT_AMD_H
#define REFRAKT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFK_OK 0
#define RFK_ERR_INVALID (-1)    /* bad argument or malformed input          */
#define RFK_ERR_PARSE (-2)      /* flame / DSL / YAML parse failure          */
#define RFK_ERR_COMPILE (-3)    /* hiprtc compile failure (log filled)       */
#define RFK_ERR_HIP (-4)        /* HIP runtime failure                       */
#define RFK_ERR_NODEVICE (-5)   /* no GPU visible                            */
#define RFK_ERR_OVERFLOW (-6)   /* a fixed-point histogram bin would wrap    */

typedef struct rfk_db rfk_db;           /* variation database                */
typedef struct rfk_flame rfk_flame;     /* host-side flame description       */
typedef struct rfk_kernel rfk_kernel;   /* compiled iterate kernel           */
typedef struct rfk_state rfk_state;     /* walkers + histogram of one render */

/* Text of the last error on this thread ("" if none). */
const char* rfk_last_error(void);
/* Library version string. */
const char* rfk_version(void);

/* ------------------------------------------------------------------ DB --
 * rfkt::initialize(flamedb&, std::string_view config_path)
 *   src/librefrakt/include/librefrakt/flame_info.h:159,
 *   src/librefrakt/src/flame_info.cpp:60-148.
 * Loads `<dir>/variations_new.yml` and `<dir>/common.yml`; entries that fail to
 * parse or type-check are skipped and counted in *n_errors (may be NULL).     */
int rfk_db_load(const char* config_dir, rfk_db** out, int* n_errors);
int rfk_db_free(rfk_db* db);
/* Number of variations, the i-th name (sorted), and an error message. */
int rfk_db_variation_count(const rfk_db* db);
const char* rfk_db_variation_name(const rfk_db* db, int i);
const char* rfk_db_error(const rfk_db* db, int i);
/* Flamedb value hash (flame_info.cpp:210-222 recalc_hash equivalent). */
uint64_t rfk_db_hash(const rfk_db* db);

/* --------------------------------------------------------------- flame --
 * rfkt::flame::deserialize (src/librefrakt/src/flame_types.cpp:437-485):
 * parses the flame JSON document (center_x, center_y, scale, rotate, gamma,
 * brightness, vibrancy, highlight_power, gamma_threshold, mod_hue, mod_sat,
 * mod_val, xforms[], final_xform, palette). `palette` is either librefrakt's
 * zlib+base64 blob of 256 (h,s,v) doubles or a JSON array of [h,s,v] triples;
 * an optional `chaos` n x n array enables xaos.                              */
int rfk_flame_from_json(const rfk_db* db, const char* json, rfk_flame** out);
/* rfkt::import_flam3 (flame_types.cpp:131-239): flam3 XML -> flame. */
int rfk_flame_from_flam3(const rfk_db* db, const char* xml, rfk_flame** out);
/* rfkt::flame::serialize (flame_types.cpp:366-397); palette written as [h,s,v]
 * triples. Writes at most cap bytes (NUL-terminated); *needed gets the size.  */
int rfk_flame_to_json(const rfk_flame* f, char* buf, size_t cap, size_t* needed);
int rfk_flame_free(rfk_flame* f);
/* rfkt::flame::size_reals (flame_types.h:602-613). */
int rfk_flame_size_reals(const rfk_flame* f, size_t* out);
/* rfkt::flame::affine_indices (flame_types.h:674-705); *count set always,
 * at most cap entries written.                                              */
int rfk_flame_affine_indices(const rfk_flame* f, size_t* out, size_t cap, size_t* count);
/* rfkt::flame::pack_samples (flame_types.h:707-712): `count` samples at
 * t = start + offset*i, each size_reals + 768 doubles (host), for a w x h view. */
int rfk_flame_pack_samples(const rfk_flame* f, double start, double offset, int count, int w, int h,
                           double* out /* host */, size_t cap);
/* Number of standard xforms and whether a final xform exists. */
int rfk_flame_xform_count(const rfk_flame* f, int* n_xforms, int* has_final);
/* Structural alignment of keyframes for interpolation (the interpolator's
 * interp_by_weight path, src/refraktor/src/refraktor.cpp:146-198): every
 * output flame holds the concatenation of all inputs' xforms, with weight 0
 * on the xforms that belong to other keyframes.                              */
int rfk_flames_align(const rfk_db* db, rfk_flame* const* in, int n, rfk_flame** out);

/* ------------------------------------------------------------ compiler --
 * rfkt::flame_compiler::get_flame_kernel / make_source
 * (src/librefrakt/include/librefrakt/flame_compiler.h:182-183,
 *  src/librefrakt/src/flame_compiler.cpp:448-656).
 * Generates gfx950 HIP source for the flame's xforms from the DSL, compiles it
 * with hiprtc (cached on disk by source hash) and loads it. On compile failure
 * returns RFK_ERR_COMPILE with the hiprtc log in `log` (may be NULL).        */
#define RFK_PRECISION_F32 0
#define RFK_PRECISION_F64 1
typedef struct rfk_exec_config {
    uint32_t threads_per_block;  /* walkers per block: 64..1024, power of 2 (0 = 256) */
    uint32_t atomic_width;       /* 32-bit bins: four u32 atomics per draw (32) or two u64 (64); 0 = 32 */
    uint32_t bin_tile_log2;      /* LDS bin cache of 2^n slots per block, 6..12 (0 = off) */
    uint32_t hist_bits;          /* histogram bin width: 32 (u32 RGBA) or 64 (u64 RGBA); 0 = 32 */
    uint32_t bin_queue;          /* 1: draws go to a record queue, sorted by 64x64 screen tile and summed
                                    per tile in LDS (no contended atomics); excludes bin_tile_log2 */
} rfk_exec_config;
int rfk_compile(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg,
                rfk_kernel** out, char* log, size_t log_cap);
/* Compile only (no device needed): runs the generator and hiprtc for gfx950
 * and reports the code-object size; used to check generated sources offline. */
int rfk_compile_check(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg,
                      size_t* code_bytes, char* log, size_t log_cap);
/* The generated HIP source (make_source); writes at most cap bytes. */
int rfk_make_source(const rfk_db* db, const rfk_flame* f, int precision, const rfk_exec_config* cfg,
                    char* buf, size_t cap, size_t* needed);
int rfk_kernel_free(rfk_kernel* k);
int rfk_kernel_size_reals(const rfk_kernel* k, size_t* out);
int rfk_kernel_threads_per_block(const rfk_kernel* k, uint32_t* out);
double rfk_kernel_compile_ms(const rfk_kernel* k);

/* --------------------------------------------------------------- render --
 * rfkt::flame_kernel::warmup (flame_compiler.h:99-100, flame_compiler.cpp:1037-1108).
 * samples: host doubles from rfk_flame_pack_samples (>= 4 samples, n_reals
 * total). Runs host fix_rotation, the Catmull-Rom coefficient kernel, sample
 * evaluation, per-sample precalc + anti-alias jitter, walker seeding and
 * `fuse` un-binned passes. `walkers` is rounded up to a whole number of blocks.
 * temporal_multiplier multiplies the number of temporal samples (blocks).
 * stream: a hipStream_t or NULL (library stream). The state owns a zeroed
 * fixed-point RGBA u32 histogram of width x height.                         */
int rfk_warmup(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t width, uint32_t height,
               uint32_t seed, uint32_t fuse, uint32_t walkers, int temporal_multiplier, rfk_state** out);
/* rfk_warmup with an explicit stream id: the Philox key of every draw of this
 * render is (seed, stream_id); stream f of a batch reproduces frame f.       */
int rfk_warmup_ex(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t width, uint32_t height,
                  uint32_t seed, uint32_t stream_id, uint32_t fuse, uint32_t walkers, int temporal_multiplier,
                  rfk_state** out);
/* Re-initialise an existing state of the same kernel for a new frame with
 * its allocations kept (the reference allocates a new state per warmup):
 * clears the histogram and walker passes, evaluates `samples`, runs precalc,
 * seeds walkers and the fuse passes. Same image size and walker count.     */
int rfk_rewarm(rfk_kernel* k, rfk_state* s, const double* samples, size_t n_reals, uint32_t seed, uint32_t stream_id,
               uint32_t fuse);
int rfk_state_free(rfk_state* s);

/* ---------------------------------------------------------------- jpeg --
 * rfkt::jpeg_encoder::encode_image (interface/jpeg_encoder.h:11-24; the nvJPEG
 * backend, impl/jpeg_encoder/nvjpeg.cpp:124-196, and the stb host backend,
 * cpu_jpeg_encoder.cpp:14-41): an 8-bit RGB (or RGBA, alpha ignored) device
 * image to a baseline JFIF file, 4:4:4 YCbCr, the JPEG standard's Annex K
 * tables scaled by the IJG quality rule (1..100; above 100 = 100), a restart
 * marker every 8 MCUs. rfk_jpeg_encode only enqueues on `stream` (like
 * nvjpegEncodeImage); rfk_jpeg_retrieve waits for it and copies the file
 * (out = NULL: *size only, like nvjpegEncodeRetrieveBitstream). One encode
 * in flight per encoder: a second rfk_jpeg_encode waits for the first.     */
typedef struct rfk_jpeg rfk_jpeg;
int rfk_jpeg_create(rfk_jpeg** out);
int rfk_jpeg_encode(rfk_jpeg* j, const void* image /* device */, int in_format /* RFK_PIXEL_RGB8 / RGBA8 */,
                    size_t pitch_bytes, uint32_t width, uint32_t height, int quality, void* stream);
int rfk_jpeg_retrieve(rfk_jpeg* j, uint8_t* out /* host */, size_t cap, size_t* size);
int rfk_jpeg_free(rfk_jpeg* j);

/* ----------------------------------------------------------- autotune --
 * No reference counterpart (librefrakt fixes its launch shape per build).
 * Compiles the flame once per candidate shape, renders probe_iters after a
 * quarter-size untimed bin, and records iterations/s of the device time in
 * each candidate; *best is the index of the fastest. Candidates that fail to
 * compile or launch keep their error code in `status`.                     */
typedef struct rfk_tune_candidate {
    uint32_t threads_per_block, atomic_width, bin_tile_log2, walkers; /* in  */
    uint32_t bin_queue;                                               /* in  */
    double iters_per_s, compile_ms;                                   /* out */
    int status;                                                       /* out */
} rfk_tune_candidate;
int rfk_autotune(const rfk_db* db, const rfk_flame* f, int precision, const double* samples, size_t n_reals,
                 uint32_t width, uint32_t height, uint64_t probe_iters, rfk_tune_candidate* cands, size_t n_cands,
                 size_t* best);

/* ---------------------------------------------------------- animation --
 * No single reference counterpart: librefrakt renders an animation one frame
 * at a time (src/refrakt-gui/src/gui/modals/render_modal.cpp:262-364,
 * warmup + bin per frame). A batch renders `frames` frames of the keyframe
 * samples (>= 4 packed samples, Catmull-Rom segments between the middle ones)
 * in ONE persistent launch: every frame has its own histogram and iteration
 * budget, work items (frame, block) are pulled longest-frame-first from a
 * device queue. With a kernel compiled with bin_queue the batch is instead a
 * sequence of chunk launches: each runs the next slice of passes of up to
 * 2048 / (walker blocks per frame) frames and sorts and sums the chunk's
 * records into those frames' histograms (the same histograms, bit for bit).
 * rfk_batch_run can capture the whole sequence in a hipGraph on its first
 * call and replay it afterwards. rfk_batch_info: grid = workgroups of the
 * (largest) launch, items = (frame, block) items, or (slice, block) items
 * with bin_queue.                                                           */
typedef struct rfk_batch rfk_batch;
typedef struct rfk_batch_frame_info {
    uint32_t* histogram;      /* device, height*width*4 bins of hist_bits     */
    double quality;           /* mean density (draw units)                    */
    uint64_t total_draws, total_passes, bad_values;
    uint32_t passes_per_walker, walkers;
    uint32_t hist_bits;       /* 32 or 64                                     */
    /* what renders the frame: its temporal samples' parameter rows (device,
     * `blocks` rows of param_stride reals of the kernel's precision, after
     * precalc; a row's first size_reals are the flame layout, the reals
     * after them the kernel's own xform-choice table) and palettes (device,
     * blocks x 256 RGBA8); its walkers run the streams (seed, stream_id)
     * with `fuse` passes before binning                                     */
    const void* params;
    const uint32_t* palettes;
    uint32_t param_stride, blocks, seed, stream_id, fuse;
} rfk_batch_frame_info;
int rfk_batch_create(rfk_kernel* k, void* stream, const double* samples, size_t n_reals, uint32_t frames,
                     const uint64_t* iters_per_frame /* host */, uint32_t width, uint32_t height,
                     uint32_t walkers_per_frame, uint32_t seed, uint32_t fuse, rfk_batch** out);
int rfk_batch_run(rfk_batch* b, int use_graph, double* kernel_ms);
int rfk_batch_frame(const rfk_batch* b, uint32_t frame, rfk_batch_frame_info* out);
int rfk_batch_info(const rfk_batch* b, uint32_t* frames, uint32_t* width, uint32_t* height, uint32_t* grid,
                   uint32_t* items);
int rfk_batch_free(rfk_batch* b);

/* rfkt::flame_kernel::bailout_args (flame_compiler.h:92-96). iters is an exact
 * iteration budget (rounded up to whole passes of all walkers; 0 = no budget),
 * millis a device-clock time limit (0 = none), quality a target mean density
 * in draws per pixel (0 = none). At least one must be set.                  */
typedef struct rfk_bailout {
    uint64_t iters;
    double millis;
    double quality;
} rfk_bailout;

/* rfkt::flame_kernel::bin_result (flame_compiler.h:27-35) plus counters. */
typedef struct rfk_bin_result {
    double quality;           /* accumulated mean density (draw units) this call */
    double elapsed_ms;        /* device clock, launch to end                 */
    double kernel_ms;         /* hipEvent time of the iterate launch(es)     */
    uint64_t total_passes;    /* iterations (walker steps) done this call    */
    uint64_t total_draws;     /* in-bounds iterations with opacity > 0       */
    uint64_t total_bins;      /* width * height                              */
    double passes_per_thread;
    double max_density;       /* max bin density after the call (draw units) */
    uint64_t bad_values;      /* walkers reset after a non-finite / huge value */
    uint64_t hit_sum;         /* sum of 255*opacity over draws (quality units) */
} rfk_bin_result;

/* rfkt::flame_kernel::bin (flame_compiler.cpp:866-955); synchronous. */
int rfk_bin(rfk_kernel* k, rfk_state* s, const rfk_bailout* bo, rfk_bin_result* out);
/* Asynchronous form: enqueues the work on the state's stream and returns;
 * rfk_bin_wait collects the result.                                         */
int rfk_bin_async(rfk_kernel* k, rfk_state* s, const rfk_bailout* bo);
int rfk_bin_wait(rfk_state* s, rfk_bin_result* out);
/* rfkt::flame_kernel::saved_state::abort_binning (flame_compiler.h:85-88).
 * Thread-safe; the running bin launch stops at its next check.             */
int rfk_abort(rfk_state* s);

/* Oracle entry (no reference counterpart): bins a device list of n points
 * (x, y, color, opacity) given in plane space, through the state's sample-0
 * screen transform and palette, with the same draw code as the iterate kernel. */
int rfk_bin_points(rfk_state* s, const float* xyco /* device, n*4 */, size_t n);
/* Test entry: runs `passes` passes and records every walker's (x, y, color,
 * opacity) after each pass and every wave's xform index (device buffers of
 * passes*walkers*4 floats and passes*(walkers/64) u32).                     */
int rfk_trace(rfk_kernel* k, rfk_state* s, uint32_t passes, float* trace, uint32_t* xform_ids);

/* State accessors (device pointers remain owned by the state). */
typedef struct rfk_state_info {
    uint32_t width, height;
    uint32_t walkers, threads_per_block, blocks;
    uint32_t size_reals, param_stride;
    uint32_t seed;
    uint64_t passes_done;     /* passes per walker completed (block 0)       */
    uint32_t* histogram;      /* device, height*width*4 bins (R,G,B,W) of hist_bits */
    float* params;            /* device, blocks*param_stride floats          */
    uint32_t* palettes;       /* device, blocks*256 packed RGBA8             */
    float* walker_x;          /* device, walkers                             */
    float* walker_y;
    float* walker_c;
    void* stream;             /* hipStream_t                                 */
    double quality;           /* accumulated mean density over all bin calls */
    uint32_t hist_bits;       /* 32 or 64                                    */
    uint32_t* walker_last;    /* device, walkers: last xform (xaos flames), else NULL */
} rfk_state_info;
int rfk_state_get_info(const rfk_state* s, rfk_state_info* out);
/* Zero the histogram and counters (walkers keep their positions). */
int rfk_state_clear_histogram(rfk_state* s);

/* -------------------------------------------------------- image kernels --
 * Density statistics: calculate_histogram<4> (assets/kernels/density_histo.cu:1-41)
 * over the fixed-point histogram: 97 log2 bins (granularity 4) and max density. */
int rfk_density_histogram(rfk_state* s, uint64_t* bins97 /* host, 97 */, double* max_density);

/* Density estimation (no reference counterpart: flam3-style variable-width
 * filter). Writes a float4 RGBA image in draw units (device, w*h*4 floats).
 * Radius h = max_radius / n^curve clamped to >= min_radius, n = bin density;
 * kernel exp(-2 (r/h)^2), r <= h, normalised per source bin. Enqueued on the
 * state's stream: the image is complete once that stream has drained
 * (rfk_synchronize, or any later work on the stream).                       */
int rfk_density_estimate(rfk_state* s, float max_radius, float min_radius, float curve, float* out_rgba);

/* rfkt::tonemapper::run (src/librefrakt/include/librefrakt/image/tonemapper.h:18-22,
 * src/librefrakt/src/image/tonemapper.cpp:37-56; assets/kernels/tonemap.cu:50-105). */
#define RFK_PIXEL_FLOAT4 0
#define RFK_PIXEL_HALF4 1
#define RFK_PIXEL_FLOAT3 2
#define RFK_PIXEL_HALF3 3
#define RFK_PIXEL_RGBA8 4
#define RFK_PIXEL_RGB8 5 /* conversion output only */
typedef struct rfk_tonemap_args {
    double quality;      /* mean density (state quality); <= 0 uses the state's */
    double gamma;
    double brightness;
    double vibrancy;
    double max_density;  /* informational, as in the reference                */
} rfk_tonemap_args;
/* Tonemap the state's fixed-point histogram, on the state's stream. Like
 * tonemapper::run the call only enqueues the work.                          */
int rfk_tonemap(rfk_state* s, void* out /* device */, int pixel_format, const rfk_tonemap_args* a);
/* density_hdr (assets/kernels/tonemap.cu:159-233): remap a
 * tonemapped float4 / half4 image (device) by the CDF of the render's
 * 97-bin density histogram (host, from rfk_density_histogram) into float4. */
int rfk_density_hdr(const void* tonemapped, int pixel_format, size_t npix, const uint64_t* bins97 /* host */,
                    float* out_rgba, void* stream);
/* Tonemap a float4 RGBA image in draw units (e.g. the density-estimated
 * one), enqueued on `stream`.                                               */
int rfk_tonemap_image(const float* in_rgba, uint32_t width, uint32_t height, void* out, int pixel_format,
                      const rfk_tonemap_args* a, void* stream);

/* rfkt::converter (src/librefrakt/include/librefrakt/image/converter.h:14-109,
 * assets/kernels/convert.cu:72-98): convert a tonemapped HALF3 / HALF4 image
 * (device) into FLOAT4, HALF4, FLOAT3, HALF3, RGBA8 or RGB8 (device),
 * enqueued on `stream`. 8-bit channels are min(v * 255, 255) truncated (0
 * for negative v); a 3-channel input gets alpha 1. flags: RFK_CONVERT_FLIP
 * writes row y to row height-1-y, RFK_CONVERT_SWAP exchanges R and B. (The
 * reference's to_float3 for half4 inputs launches its half3 kernel; here
 * every pair converts as its formats say.)                                  */
#define RFK_CONVERT_FLIP 1
#define RFK_CONVERT_SWAP 2
int rfk_convert_image(const void* in, int in_format, void* out, int out_format, uint32_t width, uint32_t height,
                      int flags, void* stream);

/* Catmull-Rom coefficients: generate_sample_coefficients
 * (assets/kernels/catmull.cu:33-42): device samples [nseg+3][sample_size]
 * -> segments [nseg][sample_size][4] (a,b,c,d), tension 0, alpha 0.5.       */
int rfk_catmull_coefficients(const double* samples, uint32_t sample_size, uint32_t nseg, double* segments,
                             void* stream);
/* Host fix_rotation (flame_compiler.cpp:957-1035) over `count` samples. */
int rfk_fix_rotation(double* samples /* host */, size_t sample_size, size_t flame_size, const size_t* affine_indices,
                     size_t n_indices, int count);

/* ------------------------------------------------------------ denoiser --
 * rfkt::denoiser (src/librefrakt/include/librefrakt/interface/denoiser.h:21-39,
 * factory.h:52): the reference wraps OIDN / OptiX binaries; this one is a
 * 4-level U-Net (3x3 convs, 32/64/96/128 channels, BF16 on the matrix cores,
 * FP32 accumulation) that predicts an RGB residual over the tonemapped
 * low-iteration frame, trained on the GPU by the same library.
 * Shapes: N images of H x W, H and W multiples of 8; features are BF16 NHWC
 * with 8 channels, base / target / output are FP32 RGB (N*H*W*3).        */
typedef struct rfk_denoiser rfk_denoiser;
int rfk_denoiser_create(uint64_t seed, rfk_denoiser** out);
int rfk_denoiser_free(rfk_denoiser* d);
int rfk_denoiser_param_count(const rfk_denoiser* d, size_t* weights, size_t* biases);
int rfk_denoiser_save(const rfk_denoiser* d, const char* path);
int rfk_denoiser_load(rfk_denoiser* d, const char* path);
int rfk_denoiser_get_params(const rfk_denoiser* d, float* w /* host */, float* b /* host */);
int rfk_denoiser_set_params(rfk_denoiser* d, const float* w /* host */, const float* b /* host */);
int rfk_denoiser_get_grads(const rfk_denoiser* d, float* gw /* host */, float* gb /* host */);
/* features + base image from a fixed-point histogram (hist_bits 32 or 64)
 * and its float4 tonemap */
int rfk_denoiser_features(const uint32_t* hist, int hist_bits, const float* tonemapped_rgba, size_t npix, void* feat,
                          float* base, void* stream);
/* rfk_tonemap into float4 followed by rfk_denoiser_features, in one pass
 * on the state's stream: the same feat (8 BF16 channels per pixel) and base
 * bit for bit, without the tonemapped image. a->quality <= 0: the quality of
 * the last bin call enqueued on the state, read from its counters on the
 * device, so the call may follow rfk_bin_async without rfk_bin_wait      */
int rfk_tonemap_features(rfk_state* s, const rfk_tonemap_args* a, void* feat, float* base);
/* out_rgb = clamp(base + U-Net(feat), 0, 1) */
int rfk_denoiser_forward(rfk_denoiser* d, const void* feat, const float* base, int N, int H, int W, float* out_rgb,
                         void* stream);
/* rfkt::denoiser::denoise (interface/denoiser.h:35-38): in = a tonemapped
 * image (RFK_PIXEL_FLOAT4 / HALF4 / FLOAT3 / HALF3, device, width x height,
 * any size), out = the denoised image in out_format (same set, device, must
 * not alias in), enqueued on `stream` (the reference returns a future; wait
 * on the stream or record an event). The network's density features come
 * from the tonemap's alpha, sat(log2(w) / 24) (tonemap.cu:95-98): a 4-channel
 * input gives the denoiser everything the histogram path gives it; a
 * 3-channel input has no alpha and is denoised from colour and coverage
 * alone. Output alpha: the input's (1 for a 3-channel input).             */
int rfk_denoise(rfk_denoiser* d, const void* in, int in_format, void* out, int out_format, uint32_t width,
                uint32_t height, void* stream);
/* one Adam step on mean squared error vs target; *loss = the batch MSE */
int rfk_denoiser_train_step(rfk_denoiser* d, const void* feat, const float* base, const float* target, int N, int H,
                            int W, float lr, float* loss, void* stream);
/* single layers (gradient checks): y = act(conv3x3(concat(x1, x2)) + bias),
 * dw = dL/dW for dL/dy = dy, dx = dL/dx. act: 0 none, 1 ReLU.             */
int rfk_nn_conv3x3_forward(const void* x1, int c1, const void* x2, int c2, int N, int H, int W, const void* wgt,
                           const float* bias, int Cout, int act, void* y, float* yf, void* stream);
/* as rfk_nn_conv3x3_forward with x1 read through a nearest 2x upsample
 * (x1_mode 1: x1 is H/2 x W/2) or a 2x2 average pool (2: x1 is 2H x 2W)  */
int rfk_nn_conv3x3_forward_ex(const void* x1, int c1, int x1_mode, const void* x2, int c2, int N, int H, int W,
                              const void* wgt, const float* bias, int Cout, int act, void* y, float* yf, void* stream);
/* as rfk_nn_conv3x3_forward_ex with BF16 y, and ypool = its 2x2 average
 * pool (N x H/2 x W/2 x Cout BF16; even H and W, Cout a multiple of 8), as
 * rfk_nn_avgpool2 computes it from y: the encoder's convolution + pool      */
int rfk_nn_conv3x3_forward_pool(const void* x1, int c1, int x1_mode, const void* x2, int c2, int N, int H, int W,
                                const void* wgt, const float* bias, int Cout, int act, void* y, void* ypool,
                                void* stream);
/* two full-resolution layers in one pass, A = act_a(conv3x3(x) + ba) (x:
 * cin <= 32 channels, A: 32, never stored), then y = act_b(conv3x3(A) + bb)
 * (BF16, cout_b = 32, ypool its 2x2 average pool or null) or, with out_rgb
 * instead of y, out_rgb = clamp(base + act_b(conv3x3(A) + bb)[:3], 0, 1)
 * (3 <= cout_b <= 8): equal to the layers run one by one                  */
int rfk_nn_conv3x3_pair_forward(const void* x, int cin, int N, int H, int W, const void* wa, const float* ba,
                                int act_a, const void* wb, const float* bb, int cout_b, int act_b, void* y,
                                void* ypool, const float* base, float* out_rgb, void* stream);
/* y = 2x2 average pool of x (N x H x W x C BF16, C a multiple of 8)       */
int rfk_nn_avgpool2(const void* x, int N, int H, int W, int C, void* y, void* stream);
int rfk_nn_conv3x3_wgrad(const void* x1, int c1, const void* x2, int c2, const void* dy, int Cout, int N, int H,
                         int W, float* dw, void* stream);
int rfk_nn_conv3x3_dgrad(const void* dy, int Cout, int N, int H, int W, const void* wgt, int Cin, void* dx,
                         void* stream);
/* training crops: n crops of ch x cw pixels starting at (image, y, x) triples
 * (host array) from a device stack of n_images H x W images of pixel_bytes
 * each; an image index outside [0, n_images) or a crop outside the image is
 * refused before anything is launched                                       */
int rfk_nn_gather_crops(const void* src, int n_images, int H, int W, int pixel_bytes, const int* offsets, int n,
                        int ch, int cw, void* dst, void* stream);
/* No counterpart in the reference: rfk_nn_gather_crops for callers that
 * cannot allocate device memory (tight training loops, code under graph
 * capture, harnesses that forbid hipMalloc). The crop offsets are staged in a
 * static device array of RFK_NOALLOC_MAX_CROPS triples (48 KiB, held by each
 * GPU the library runs on until the process ends) instead of a per-call
 * allocation; more crops are refused (RFK_ERR_INVALID). Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from other threads
 * or on other streams) race. Returns when the crops are in dst.             */
#define RFK_NOALLOC_MAX_CROPS 4096
int rfk_nn_gather_crops_noalloc(const void* src, int n_images, int H, int W, int pixel_bytes, const int* offsets,
                                int n, int ch, int cw, void* dst, void* stream);
/* error-path tests: the next `count` activation allocations of any denoiser
 * fail as hipMalloc failing would (RFK_ERR_HIP, nothing launched)           */
int rfk_nn_inject_alloc_failures(int count);

/* Microbenchmark of uniform-random RGBA u32 atomic increments into a buffer of
 * `footprint_bytes`; returns increments per second (the binning roofline).  */
int rfk_atomic_peak(uint64_t footprint_bytes, uint64_t increments, double* increments_per_s);

/* Device helpers for hosts without a GPU framework. */
int rfk_device_count(int* n);
int rfk_malloc(void** p, size_t bytes);
int rfk_free(void* p);
int rfk_memcpy_htod(void* dst, const void* src, size_t bytes);
int rfk_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int rfk_memcpy_dtod(void* dst, const void* src, size_t bytes);
int rfk_memset(void* dst, int value, size_t bytes);
int rfk_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* REFRAKT_AMD_H */
