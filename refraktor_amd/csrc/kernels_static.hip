// Kernels compiled ahead of time for gfx950. All are streaming kernels over
// HBM-resident images except the per-sample evaluation.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "device/flame_device.h"
#include "static_kernels.hpp"

namespace rfk::sk {

namespace {

constexpr unsigned kBlock = 256;

// --------------------------------------------------------------- Catmull-Rom
// Centripetal-style knot spacing pow(hypot(dp, 1), alpha) with tension 0 and
// alpha 0.5; coefficients of a t^3 + b t^2 + c t + d over [p1, p2].
__device__ void catmull_segment(double p0, double p1, double p2, double p3, double tension, double alpha,
                                double* out) {
#pragma clang fp contract(off)
    double t01 = pow(hypot(p1 - p0, 1.0), alpha);
    double t12 = pow(hypot(p2 - p1, 1.0), alpha);
    double t23 = pow(hypot(p3 - p2, 1.0), alpha);
    double m1 = (1.0 - tension) * (p2 - p1 + t12 * ((p1 - p0) / t01 - (p2 - p0) / (t01 + t12)));
    double m2 = (1.0 - tension) * (p2 - p1 + t12 * ((p3 - p2) / t23 - (p3 - p1) / (t12 + t23)));
    out[0] = 2.0 * (p1 - p2) + m1 + m2;
    out[1] = -3.0 * (p1 - p2) - m1 - m1 - m2;
    out[2] = m1;
    out[3] = p1;
}

__global__ void catmull_kernel(const double* __restrict__ samples, uint32_t sample_size, uint32_t nseg,
                               double* __restrict__ segments) {
#pragma clang fp contract(off)
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= uint64_t(sample_size) * nseg) return;
    const uint64_t idx = g;  // = seg * sample_size + param
    double coef[4];
    catmull_segment(samples[idx], samples[idx + sample_size], samples[idx + 2ull * sample_size],
                    samples[idx + 3ull * sample_size], 0.0, 0.5, coef);
    double4 o = {coef[0], coef[1], coef[2], coef[3]};
    reinterpret_cast<double4*>(segments)[idx] = o;
}

// evaluated without FMA contraction so the host restatement reproduces the bits
__device__ __forceinline__ double seg_eval(const double* seg4, double t) {
#pragma clang fp contract(off)
    return seg4[0] * t * t * t + seg4[1] * t * t + seg4[2] * t + seg4[3];
}

// The reference gives block b of an n-block grid t = (b % seg_size) / (n / nseg)
// inside segment b / seg_size (refactor.cu:266-270).
// With `times` given, sample s sits at global time times[s] in [0, nseg):
// segment floor(times[s]), local t = the fraction (animation batches).
template <typename OutT>
__global__ void eval_samples_kernel(const double* __restrict__ segments, uint32_t sample_size, uint32_t size_reals,
                                    uint32_t nseg, uint32_t nsamples, const int8_t* __restrict__ kind,
                                    OutT* __restrict__ params, uint32_t param_stride, uint32_t* __restrict__ palettes,
                                    const double* __restrict__ times) {
    const uint32_t per = size_reals + 256u;
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= uint64_t(per) * nsamples) return;
    const uint32_t s = uint32_t(g / per), i = uint32_t(g % per);
    uint32_t seg;
    double t;
    if (times) {
        const double gt = fmin(fmax(times[s], 0.0), double(nseg));
        seg = uint32_t(gt);
        if (seg >= nseg) seg = nseg - 1;
        t = gt - double(seg);
    } else {
        const uint32_t seg_size = max(nsamples / nseg, 1u);
        seg = s / seg_size;
        if (seg >= nseg) seg = nseg - 1;
        t = double(s % seg_size) / double(max(nsamples / nseg, 1u));
    }
    const double* base = segments + uint64_t(seg) * sample_size * 4ull;
    if (i < size_reals) {
        const int k = kind[i];
        double v;
        if (k == 0) {
            v = seg_eval(base + 4ull * i, t);
        } else {
            // polar -> cartesian (refactor.cu:279-294): stored as
            // (angle_x, angle_y, log|x|, log|y|, c, f)
            const uint32_t a = i - uint32_t(k - 1);
            const int comp = k - 1;  // 0:a 1:d 2:b 3:e
            const int col = comp >> 1;
            const double ang = seg_eval(base + 4ull * (a + col), t);
            const double mag = exp(seg_eval(base + 4ull * (a + 2 + col), t));
            v = (comp & 1) ? mag * sin(ang) : mag * cos(ang);
        }
        params[uint64_t(s) * param_stride + i] = OutT(v);
    } else {
        const uint32_t j = i - size_reals;
        const double* pb = base + 4ull * (size_reals + 3u * j);
        const double h = seg_eval(pb, t), sat = seg_eval(pb + 4, t), val = seg_eval(pb + 8, t);
        // HSV -> RGB of refactor.cu:128-140, evaluated in double, uncontracted
        auto f = [&](double n) {
#pragma clang fp contract(off)
            double k = fmod(n + h / 60.0, 6.0);
            return val - val * sat * fmax(0.0, fmin(k, fmin(4.0 - k, 1.0)));
        };
        auto to8 = [](double c) -> uint32_t {
            double v = c * 255.0;
            if (!(v > 0.0)) return 0u;
            if (v >= 255.0) return 255u;
            return uint32_t(v);
        };
        palettes[uint64_t(s) * 256u + j] = to8(f(5.0)) | (to8(f(3.0)) << 8) | (to8(f(1.0)) << 16) | (255u << 24);
    }
}

// ------------------------------------------------------------------ points
template <class HistT>
__global__ void bin_points_kernel(const float* __restrict__ screen6, const uint32_t* __restrict__ palette,
                                  const float4* __restrict__ pts, size_t n, HistT* __restrict__ hist,
                                  uint32_t width, uint32_t height) {
    __shared__ uint32_t s_pal[256];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) s_pal[i] = palette[i];
    float S[6];
    for (int k = 0; k < 6; k++) S[k] = screen6[k];
    __syncthreads();
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float4 p = pts[g];
    rfk_draw_t d = rfk_make_draw(S, s_pal, width, height, p.x, p.y, p.z, p.w);
    if (d.hit) rfk_commit_draw(hist, d);
}

// ------------------------------------------------------------ bin readers
// Histogram elements: u32 RGBA (uint4) or u64 RGBA (bin64). Draw units are
// value / 255.
struct __align__(32) bin64 {
    unsigned long long x, y, z, w;
};
__device__ __forceinline__ float4 bin_units(const uint4& h) {
    const float k = 1.0f / 255.0f;
    return {float(h.x) * k, float(h.y) * k, float(h.z) * k, float(h.w) * k};
}
__device__ __forceinline__ float4 bin_units(const bin64& h) {
    const float k = 1.0f / 255.0f;
    return {float(h.x) * k, float(h.y) * k, float(h.z) * k, float(h.w) * k};
}

// ------------------------------------------------------- density histogram
constexpr unsigned kHistoBlock = 1024;
template <class BinT>
__global__ __launch_bounds__(kHistoBlock) void density_histogram_kernel(const BinT* __restrict__ hist, size_t npix,
                                                                        unsigned long long* __restrict__ bins97,
                                                                        unsigned long long* __restrict__ max_density) {
    constexpr unsigned kBins = 4 * 24 + 1, kWaves = kHistoBlock / 64;
    // a workgroup counts fewer than 2^32 pixels: 32-bit LDS counters, one
    // set per wave (the lanes of one wave that share a bin serialise on it,
    // the waves of the workgroup no longer do), summed and added into the
    // 64-bit global bins once
    __shared__ unsigned s_wbins[kWaves][kBins];
    __shared__ unsigned long long s_max;
    for (unsigned i = threadIdx.x; i < kWaves * kBins; i += blockDim.x) (&s_wbins[0][0])[i] = 0;
    unsigned* s_bins = s_wbins[threadIdx.x >> 6];
    if (threadIdx.x == 0) s_max = 0;
    __syncthreads();
    const float log_div = 4.0f / logf(2.0f);
    // bin 0 (the empty and faint pixels, most of an image) and the maximum
    // are kept per thread and added once: one LDS address for every lane
    // would serialise the atomics
    unsigned long long low = 0, mx = 0;
    // four pixels' loads in flight per thread before any is used
    constexpr unsigned U = 4;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x; g < npix; g += U * stride) {
        float v[U];
#pragma unroll
        for (unsigned u = 0; u < U; u++) {
            const size_t gi = g + u * stride;
            v[u] = gi < npix ? bin_units(hist[gi]).w : 0.0f;
        }
#pragma unroll
        for (unsigned u = 0; u < U; u++) {
            if (g + u * stride >= npix) break;
            const float value = v[u];
            if (value >= 1.0f) {
                unsigned long long b = (unsigned long long)(logf(value) * log_div) + 1ull;
                if (b >= kBins) b = kBins - 1;
                atomicAdd(&s_bins[b], 1u);
                mx = max(mx, (unsigned long long)value);
            } else {
                low++;
            }
        }
    }
    if (low) atomicAdd(&s_bins[0], unsigned(low));
    if (mx) atomicMax(&s_max, mx);
    __syncthreads();
    for (unsigned i = threadIdx.x; i < kBins; i += blockDim.x) {
        unsigned long long sum = 0;
#pragma unroll
        for (unsigned w = 0; w < kWaves; w++) sum += s_wbins[w][i];
        if (sum) atomicAdd(&bins97[i], sum);
    }
    if (threadIdx.x == 0) atomicMax(max_density, s_max);
}

// ------------------------------------------------------ density estimation
// Level of a source bin: h = max(min_r, max_r * max(n,1)^-curve), quantised
// to 1/16 pixel; -1 for an empty bin.
template <class BinT>
__global__ void de_levels_kernel(const BinT* __restrict__ hist, size_t npix, float max_radius, float min_radius,
                                 float curve, int max_level, int* __restrict__ levels) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= npix) return;
    if (hist[g].w == 0u) {
        levels[g] = -1;
        return;
    }
    const float n = bin_units(hist[g]).w;
    float h = max_radius * powf(fmaxf(n, 1.0f), -curve);
    h = fmaxf(h, min_radius);
    int lvl = int(h * 16.0f + 0.5f);
    levels[g] = min(lvl, max_level);
}

template <int TILE, class BinT>
__global__ void de_gather_kernel(const BinT* __restrict__ hist, const int* __restrict__ levels, uint32_t width,
                                 uint32_t height, int radius, const float* __restrict__ inv_z, int nlevels,
                                 float4* __restrict__ out) {
    extern __shared__ unsigned char smem[];
    const int span = TILE + 2 * radius;
    float4* s_col = reinterpret_cast<float4*>(smem);
    int* s_lvl = reinterpret_cast<int*>(s_col + span * span);
    float* s_iz = reinterpret_cast<float*>(s_lvl + span * span);
    __shared__ int s_maxlv;
    const int x0 = int(blockIdx.x) * TILE - radius, y0 = int(blockIdx.y) * TILE - radius;
    const int tid = int(threadIdx.y) * TILE + int(threadIdx.x);
    if (tid == 0) s_maxlv = -1;
    for (int i = tid; i < nlevels; i += TILE * TILE) s_iz[i] = inv_z[i];
    __syncthreads();
    int maxlv = -1;
    for (int i = tid; i < span * span; i += TILE * TILE) {
        const int xx = x0 + i % span, yy = y0 + i / span;
        int lv = -1;
        float4 c = {0.f, 0.f, 0.f, 0.f};
        if (xx >= 0 && yy >= 0 && xx < int(width) && yy < int(height)) {
            const size_t p = size_t(yy) * width + size_t(xx);
            lv = levels[p];
            if (lv >= 0) c = bin_units(hist[p]);
        }
        s_lvl[i] = lv;
        s_col[i] = c;
        maxlv = max(maxlv, lv);
    }
    atomicMax(&s_maxlv, maxlv);
    __syncthreads();
    // the span's non-empty sources, compacted in span (row-major) order: a
    // contiguous range of entries per thread, counted, scanned, written
    unsigned short* s_list = reinterpret_cast<unsigned short*>(s_iz + nlevels);
    __shared__ int s_wave_tot[TILE * TILE / 64], s_count;
    {
        const int total = span * span, per = (total + TILE * TILE - 1) / (TILE * TILE);
        const int i0 = min(tid * per, total), i1 = min(i0 + per, total);
        int cnt = 0;
        for (int i = i0; i < i1; i++) cnt += s_lvl[i] >= 0;
        int incl = cnt;
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(incl, d);
            if (__lane_id() >= unsigned(d)) incl += u;
        }
        if (__lane_id() == 63) s_wave_tot[tid >> 6] = incl;
        __syncthreads();
        int base = 0, all = 0;
        for (int w = 0; w < TILE * TILE / 64; w++) {
            if (w < (tid >> 6)) base += s_wave_tot[w];
            all += s_wave_tot[w];
        }
        int k = base + incl - cnt;
        for (int i = i0; i < i1; i++)
            if (s_lvl[i] >= 0) s_list[k++] = (unsigned short)((i % span) | ((i / span) << 8));
        if (tid == 0) s_count = all;
        __syncthreads();
    }
    const int ox = int(blockIdx.x) * TILE + int(threadIdx.x), oy = int(blockIdx.y) * TILE + int(threadIdx.y);
    if (ox >= int(width) || oy >= int(height)) return;
    // no source of this span reaches further than its widest kernel
    const int reach = min(radius, (s_maxlv + 15) / 16);
    float4 acc = {0.f, 0.f, 0.f, 0.f};
    if (s_count <= (2 * reach + 1) * (2 * reach + 1)) {
        // fewer sources than window offsets: visit the sources instead, in
        // the same (row-major) order as the window loop below, with the same
        // arithmetic, so both paths give the same bits
        const int cx = int(threadIdx.x) + radius, cy = int(threadIdx.y) + radius;
        for (int i = 0; i < s_count; i++) {
            const int e = s_list[i];
            const int sx = e & 0xFF, sy = e >> 8;
            const int dx = sx - cx, dy = sy - cy;
            const int si = sy * span + sx;
            const int lv = s_lvl[si];
            const float h = float(lv) * (1.0f / 16.0f);
            const float d2 = float(dx * dx + dy * dy);
            if (d2 > h * h) continue;
            const float wgt = (d2 == 0.0f ? 1.0f : __expf(-2.0f * d2 / (h * h))) * s_iz[lv];
            const float4 c = s_col[si];
            acc.x = fmaf(c.x, wgt, acc.x);
            acc.y = fmaf(c.y, wgt, acc.y);
            acc.z = fmaf(c.z, wgt, acc.z);
            acc.w = fmaf(c.w, wgt, acc.w);
        }
        out[size_t(oy) * width + size_t(ox)] = acc;
        return;
    }
    for (int dy = -reach; dy <= reach; dy++) {
        const int row = (int(threadIdx.y) + radius + dy) * span + int(threadIdx.x) + radius;
        for (int dx = -reach; dx <= reach; dx++) {
            const int lv = s_lvl[row + dx];
            if (lv < 0) continue;
            const float h = float(lv) * (1.0f / 16.0f);
            const float d2 = float(dx * dx + dy * dy);
            if (d2 > h * h) continue;
            const float wgt = (d2 == 0.0f ? 1.0f : __expf(-2.0f * d2 / (h * h))) * s_iz[lv];
            const float4 c = s_col[row + dx];
            acc.x = fmaf(c.x, wgt, acc.x);
            acc.y = fmaf(c.y, wgt, acc.y);
            acc.z = fmaf(c.z, wgt, acc.z);
            acc.w = fmaf(c.w, wgt, acc.w);
        }
    }
    out[size_t(oy) * width + size_t(ox)] = acc;
}

// ------------------------------------------------------------------ tonemap
struct half3_t {
    __half x, y, z;
};
struct half4_t {
    __half x, y, z, w;
};

__device__ __forceinline__ float sat(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
// x^y for x >= 0 from the hardware's base-2 exponential and logarithm
// (v_exp_f32, v_log_f32: about 1 ulp each): the reference builds its
// tonemapper with fast math (tonemapper.cpp:20), where powf is this; the
// library powf's exact rounding made the tonemap VALU-bound
__device__ __forceinline__ float fast_pow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// One pixel of the log-density tonemap (tonemap.cu:50-105): draw units in,
// RGBA out (zero for an empty bin)
__device__ __forceinline__ float4 tonemap_pixel(float4 col, float scale_constant, float brightness, float gamma,
                                                float vibrancy) {
    float4 o = {0.f, 0.f, 0.f, 0.f};
    if (col.w != 0.0f) {
        const float input_density = col.w;
        const float factor = 0.5f * brightness * logf(1.0f + col.w * scale_constant) * 0.434294481903251827651128918916f / col.w;
        col.x *= factor;
        col.y *= factor;
        col.z *= factor;
        col.w *= factor;
        // tonemap.cu:88-94 takes z = w^(1/gamma) and the colour products in
        // double; FP32 here (a few ulp of the FP32 result, far inside the
        // tonemap's tolerance against the double oracle)
        const float ig = float(1.0 / double(gamma));
        const float gamma_factor = fast_pow(col.w, ig) / col.w;
        o.x = sat(lerpf(fast_pow(col.x, ig), col.x * gamma_factor, vibrancy));
        o.y = sat(lerpf(fast_pow(col.y, ig), col.y * gamma_factor, vibrancy));
        o.z = sat(lerpf(fast_pow(col.z, ig), col.z * gamma_factor, vibrancy));
        o.w = sat(logf(input_density) * (4.0f / logf(2.0f)) / 96.0f);
    }
    return o;
}

// In: 0 float4 draw units, 1 u32 bins, 2 u64 bins
template <int In, int Fmt>
__global__ void tonemap_kernel(const void* __restrict__ in, size_t npix, void* __restrict__ out, float scale_constant,
                               float brightness, float gamma, float vibrancy) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= npix) return;
    float4 col;
    if (In == 0) col = reinterpret_cast<const float4*>(in)[g];
    else if (In == 1) col = bin_units(reinterpret_cast<const uint4*>(in)[g]);
    else col = bin_units(reinterpret_cast<const bin64*>(in)[g]);
    const float4 o = tonemap_pixel(col, scale_constant, brightness, gamma, vibrancy);
    if (Fmt == 0) {
        reinterpret_cast<float4*>(out)[g] = o;
    } else if (Fmt == 1) {
        reinterpret_cast<half4_t*>(out)[g] = {__float2half(o.x), __float2half(o.y), __float2half(o.z), __float2half(o.w)};
    } else if (Fmt == 2) {
        float* f = reinterpret_cast<float*>(out) + 3 * g;
        f[0] = o.x;
        f[1] = o.y;
        f[2] = o.z;
    } else if (Fmt == 3) {
        reinterpret_cast<half3_t*>(out)[g] = {__float2half(o.x), __float2half(o.y), __float2half(o.z)};
    } else {
        auto q = [](float v) { return uint32_t(fminf(fmaxf(v * 255.0f + 0.5f, 0.0f), 255.0f)); };
        reinterpret_cast<uint32_t*>(out)[g] = q(o.x) | (q(o.y) << 8) | (q(o.z) << 16) | (q(o.w) << 24);
    }
}

// The denoiser's input straight from the histogram, without the tonemapped
// image in between: the tonemap of tonemap_kernel and the features of the
// denoiser's features kernel from it ([R, G, B, alpha] of the tonemap,
// log2(1 + density) / 16, coverage, two zero channels; BF16) and the FP32
// base image (the tonemap's RGB)
// With `hits` (the bin call's hit counter in device memory) the scale
// constant is computed here from the call's quality, hits / (pixels x 255),
// as the host computes it, so the caller need not wait for the call.
template <class BinT>
__global__ void tonemap_features_kernel(const BinT* __restrict__ hist, size_t npix, float scale_constant,
                                        const unsigned long long* __restrict__ hits, float brightness, float gamma,
                                        float vibrancy, __bf16* __restrict__ feat, float* __restrict__ base) {
    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= npix) return;
    if (hits) {
        const double q = double(*hits) / (double(npix) * 255.0);
        scale_constant = powf(10.0f, -log10f(float(q)) - 0.5f);
    }
    const float4 col = bin_units(hist[g]);
    const float dens = col.w;
    const float4 o = tonemap_pixel(col, scale_constant, brightness, gamma, vibrancy);
    bf16x8 f;
    f[0] = __bf16(o.x);
    f[1] = __bf16(o.y);
    f[2] = __bf16(o.z);
    f[3] = __bf16(o.w);
    f[4] = __bf16(log2f(1.0f + dens) * (1.0f / 16.0f));
    f[5] = __bf16(dens > 0.0f ? 1.0f : 0.0f);
    f[6] = __bf16(0.0f);
    f[7] = __bf16(0.0f);
    *reinterpret_cast<bf16x8*>(feat + 8 * g) = f;
    base[3 * g + 0] = o.x;
    base[3 * g + 1] = o.y;
    base[3 * g + 2] = o.z;
}

// -------------------------------------------------------------- density HDR
// CDF remap of a tonemapped image (tonemap.cu:159-233): the density
// histogram's CDF at the pixel's log-density (alpha * 96 + 1, lerped between
// bins, clamped to the last bin) scales the HSV value by (cdf / 0.5)^2;
// alpha becomes 1.
__device__ __forceinline__ float4 rgb_to_hsv_dev(float4 c) {
    const float mx = fmaxf(c.x, fmaxf(c.y, c.z)), mn = fminf(c.x, fminf(c.y, c.z)), d = mx - mn;
    float h = 0.0f;
    if (d > 0.0f) {
        if (mx == c.x) h = fmodf((c.y - c.z) / d, 6.0f);
        else if (mx == c.y) h = (c.z - c.x) / d + 2.0f;
        else h = (c.x - c.y) / d + 4.0f;
        if (h < 0.0f) h += 6.0f;
    }
    return {h, mx > 0.0f ? d / mx : 0.0f, mx, c.w};  // h in sextants [0, 6)
}

__device__ __forceinline__ float4 hsv_to_rgb_dev(float4 hsv) {
    const float h = hsv.x, s = hsv.y, v = hsv.z;
    const int i = int(floorf(h)) % 6;
    const float f = h - floorf(h);
    const float p = v * (1.0f - s), q = v * (1.0f - s * f), t = v * (1.0f - s * (1.0f - f));
    switch (i) {
        case 0: return {v, t, p, hsv.w};
        case 1: return {q, v, p, hsv.w};
        case 2: return {p, v, t, hsv.w};
        case 3: return {p, q, v, hsv.w};
        case 4: return {t, p, v, hsv.w};
        default: return {v, p, q, hsv.w};
    }
}

template <bool Half>
__global__ void density_hdr_kernel(const void* __restrict__ in, size_t npix, const unsigned long long* __restrict__ bins97,
                                   float4* __restrict__ out) {
    constexpr int kBins = 97;
    __shared__ float cdf[kBins];
    if (threadIdx.x == 0) {
        double total = 0.0;
        for (int i = 0; i < kBins; i++) total += double(bins97[i]);
        double run = 0.0;
        for (int i = 0; i < kBins; i++) {
            run += double(bins97[i]);
            cdf[i] = total > 0.0 ? float(run / total) : 0.0f;
        }
    }
    __syncthreads();
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= npix) return;
    float4 c;
    if (Half) {
        const half4_t h = reinterpret_cast<const half4_t*>(in)[g];
        c = {__half2float(h.x), __half2float(h.y), __half2float(h.z), __half2float(h.w)};
    } else {
        c = reinterpret_cast<const float4*>(in)[g];
    }
    const float power = c.w * float(kBins - 1) + 1.0f;
    const int lo = min(int(power), kBins - 1), hi = min(lo + 1, kBins - 1);
    const float frac = power - floorf(power);
    const float v = lerpf(cdf[lo], cdf[hi], frac);
    float4 hsv = rgb_to_hsv_dev(c);
    const float gain = v / 0.5f;
    hsv.z *= gain * gain;
    const float4 o = hsv_to_rgb_dev(hsv);
    out[g] = {o.x, o.y, o.z, 1.0f};
}

// --------------------------------------------------------------- atomics
__global__ void atomic_peak_kernel(uint32_t* __restrict__ buf, uint64_t nbins, uint32_t iters, uint32_t seed) {
    uint32_t h = (blockIdx.x * blockDim.x + threadIdx.x) * 0x9E3779B1u ^ seed;
    for (uint32_t i = 0; i < iters; i++) {
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        const uint64_t bin = (uint64_t(h) * nbins) >> 32;
        uint32_t* p = buf + 4ull * bin;
        atomicAdd(p + 0, 1u);
        atomicAdd(p + 1, 1u);
        atomicAdd(p + 2, 1u);
        atomicAdd(p + 3, 1u);
        h += i;
    }
}

unsigned nblocks(size_t n, unsigned b = kBlock) { return unsigned((n + b - 1) / b); }

}  // namespace

hipError_t catmull(const double* samples, uint32_t sample_size, uint32_t nseg, double* segments, hipStream_t s) {
    size_t n = size_t(sample_size) * nseg;
    if (!n) return hipSuccess;
    catmull_kernel<<<nblocks(n), kBlock, 0, s>>>(samples, sample_size, nseg, segments);
    return hipGetLastError();
}

hipError_t eval_samples(const double* segments, uint32_t sample_size, uint32_t size_reals, uint32_t nseg,
                        uint32_t nsamples, const int8_t* kind, int precision, void* params, uint32_t param_stride,
                        uint32_t* palettes, hipStream_t s, const double* times) {
    size_t n = size_t(size_reals + 256u) * nsamples;
    if (precision == 1)
        eval_samples_kernel<double><<<nblocks(n), kBlock, 0, s>>>(segments, sample_size, size_reals, nseg, nsamples,
                                                                   kind, (double*)params, param_stride, palettes, times);
    else
        eval_samples_kernel<float><<<nblocks(n), kBlock, 0, s>>>(segments, sample_size, size_reals, nseg, nsamples,
                                                                  kind, (float*)params, param_stride, palettes, times);
    return hipGetLastError();
}

// p[i] = v, except w for i in [at, at + m)
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n, size_t at, size_t m, uint32_t w) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = (i - at < m) ? w : v;
}

hipError_t fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) { return fill_u32_except(p, v, n, 0, 0, v, s); }

hipError_t fill_u32_except(uint32_t* p, uint32_t v, size_t n, size_t at, size_t m, uint32_t w, hipStream_t s) {
    if (!n) return hipSuccess;
    unsigned g = std::min(nblocks(n), 8192u);
    fill_u32_kernel<<<g, kBlock, 0, s>>>(p, v, n, at, m, w);
    return hipGetLastError();
}

hipError_t bin_points(const float* screen6, const uint32_t* palette, const float* pts, size_t n, uint32_t* hist,
                      bool hist64, uint32_t width, uint32_t height, hipStream_t s) {
    if (!n) return hipSuccess;
    const float4* p4 = reinterpret_cast<const float4*>(pts);
    if (hist64)
        bin_points_kernel<<<nblocks(n), kBlock, 0, s>>>(screen6, palette, p4, n,
                                                         reinterpret_cast<unsigned long long*>(hist), width, height);
    else
        bin_points_kernel<<<nblocks(n), kBlock, 0, s>>>(screen6, palette, p4, n, hist, width, height);
    return hipGetLastError();
}

hipError_t density_histogram(const uint32_t* hist, bool hist64, size_t npix, unsigned long long* bins97,
                             unsigned long long* max_density, hipStream_t s) {
    // one workgroup per CU (every workgroup ends with one global atomic per
    // non-empty bin, on the same 98 addresses: 2048 workgroups of 256 took
    // 33.5 us at 1080p, 256 of them 19.6 us), 16 waves each
    unsigned grid = std::min(nblocks(npix, kHistoBlock), 256u);
    if (grid == 0) grid = 1;
    if (hist64)
        density_histogram_kernel<<<grid, kHistoBlock, 0, s>>>(reinterpret_cast<const bin64*>(hist), npix, bins97,
                                                              max_density);
    else
        density_histogram_kernel<<<grid, kHistoBlock, 0, s>>>(reinterpret_cast<const uint4*>(hist), npix, bins97,
                                                              max_density);
    return hipGetLastError();
}

template <class BinT>
static hipError_t density_estimate_t(const BinT* hist, uint32_t width, uint32_t height, float max_radius,
                                     float min_radius, float curve, const float* inv_z, int nlevels, int radius,
                                     int* level_scratch, float* out_rgba, hipStream_t s) {
    size_t npix = size_t(width) * height;
    de_levels_kernel<<<nblocks(npix), kBlock, 0, s>>>(hist, npix, max_radius, min_radius, curve, nlevels - 1,
                                                      level_scratch);
    constexpr int T = 16;
    const int span = T + 2 * radius;
    size_t shmem = size_t(span) * span * (sizeof(float4) + sizeof(int) + sizeof(unsigned short)) +
                   size_t(nlevels) * sizeof(float);
    dim3 grid((width + T - 1) / T, (height + T - 1) / T);
    de_gather_kernel<T><<<grid, dim3(T, T), shmem, s>>>(hist, level_scratch, width, height, radius, inv_z, nlevels,
                                                        reinterpret_cast<float4*>(out_rgba));
    return hipGetLastError();
}

hipError_t density_estimate(const uint32_t* hist, bool hist64, uint32_t width, uint32_t height, float max_radius,
                            float min_radius, float curve, const float* inv_z, int nlevels, int radius,
                            int* level_scratch, float* out_rgba, hipStream_t s) {
    if (hist64)
        return density_estimate_t(reinterpret_cast<const bin64*>(hist), width, height, max_radius, min_radius, curve,
                                  inv_z, nlevels, radius, level_scratch, out_rgba, s);
    return density_estimate_t(reinterpret_cast<const uint4*>(hist), width, height, max_radius, min_radius, curve,
                              inv_z, nlevels, radius, level_scratch, out_rgba, s);
}

hipError_t tonemap_features(const void* hist, bool hist64, size_t npix, float scale_constant,
                            const unsigned long long* hits, float brightness, float gamma, float vibrancy, void* feat,
                            float* base, hipStream_t s) {
    if (!npix) return hipSuccess;
    if (hist64)
        tonemap_features_kernel<<<nblocks(npix), kBlock, 0, s>>>(reinterpret_cast<const bin64*>(hist), npix,
                                                                 scale_constant, hits, brightness, gamma, vibrancy,
                                                                 reinterpret_cast<__bf16*>(feat), base);
    else
        tonemap_features_kernel<<<nblocks(npix), kBlock, 0, s>>>(reinterpret_cast<const uint4*>(hist), npix,
                                                                 scale_constant, hits, brightness, gamma, vibrancy,
                                                                 reinterpret_cast<__bf16*>(feat), base);
    return hipGetLastError();
}

hipError_t tonemap(const void* in, int in_kind, size_t npix, void* out, int format, float scale_constant,
                   float brightness, float gamma, float vibrancy, hipStream_t s) {
    if (!npix) return hipSuccess;
    if (in_kind < 0 || in_kind > 2) return hipErrorInvalidValue;
    unsigned g = nblocks(npix);
#define RFK_TM(F)                                                                                              \
    if (in_kind == 0) tonemap_kernel<0, F><<<g, kBlock, 0, s>>>(in, npix, out, scale_constant, brightness, gamma, vibrancy); \
    else if (in_kind == 1) tonemap_kernel<1, F><<<g, kBlock, 0, s>>>(in, npix, out, scale_constant, brightness, gamma, vibrancy); \
    else tonemap_kernel<2, F><<<g, kBlock, 0, s>>>(in, npix, out, scale_constant, brightness, gamma, vibrancy);
    switch (format) {
        case 0: RFK_TM(0) break;
        case 1: RFK_TM(1) break;
        case 2: RFK_TM(2) break;
        case 3: RFK_TM(3) break;
        case 4: RFK_TM(4) break;
        default: return hipErrorInvalidValue;
    }
#undef RFK_TM
    return hipGetLastError();
}

// ----------------------------------------------------------- conversion
// convert_to (assets/kernels/convert.cu:72-98): half3/half4 in; out element
// by element as float, half or 8-bit (min(v * 255, 255), truncated: exact in
// float for half inputs); a 3-channel input gets alpha 1; `swap` exchanges
// R and B; `flip` writes row y to row height - 1 - y.
template <int InC, int OutFmt>
__global__ void convert_kernel(const __half* __restrict__ in, void* __restrict__ out, uint32_t width, uint32_t height,
                               bool flip, bool swap) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t npix = size_t(width) * height;
    if (g >= npix) return;
    __half h[4];
    if constexpr (InC == 4) {
        const half4_t v = reinterpret_cast<const half4_t*>(in)[g];
        h[0] = v.x;
        h[1] = v.y;
        h[2] = v.z;
        h[3] = v.w;
    } else {
        const half3_t v = reinterpret_cast<const half3_t*>(in)[g];
        h[0] = v.x;
        h[1] = v.y;
        h[2] = v.z;
        h[3] = __float2half(1.0f);
    }
    if (swap) {
        const __half t = h[0];
        h[0] = h[2];
        h[2] = t;
    }
    size_t o = g;
    if (flip) {
        const size_t y = g / width, x = g % width;
        o = (size_t(height) - 1 - y) * width + x;
    }
    // (negative inputs, undefined in the reference's cast, give 0)
    auto u8 = [](__half v) { return uint32_t(fmaxf(fminf(__half2float(v) * 255.0f, 255.0f), 0.0f)); };
    if constexpr (OutFmt == 0) {
        reinterpret_cast<float4*>(out)[o] = {__half2float(h[0]), __half2float(h[1]), __half2float(h[2]), __half2float(h[3])};
    } else if constexpr (OutFmt == 1) {
        reinterpret_cast<half4_t*>(out)[o] = {h[0], h[1], h[2], h[3]};
    } else if constexpr (OutFmt == 2) {
        float* f = reinterpret_cast<float*>(out) + 3 * o;
        f[0] = __half2float(h[0]);
        f[1] = __half2float(h[1]);
        f[2] = __half2float(h[2]);
    } else if constexpr (OutFmt == 3) {
        reinterpret_cast<half3_t*>(out)[o] = {h[0], h[1], h[2]};
    } else if constexpr (OutFmt == 4) {
        reinterpret_cast<uint32_t*>(out)[o] = u8(h[0]) | (u8(h[1]) << 8) | (u8(h[2]) << 16) | (u8(h[3]) << 24);
    } else {
        uint8_t* b = reinterpret_cast<uint8_t*>(out) + 3 * o;
        b[0] = uint8_t(u8(h[0]));
        b[1] = uint8_t(u8(h[1]));
        b[2] = uint8_t(u8(h[2]));
    }
}

hipError_t convert_image(const void* in, int in_channels, void* out, int out_format, uint32_t width, uint32_t height,
                         bool flip, bool swap, hipStream_t s) {
    const size_t npix = size_t(width) * height;
    if (!npix) return hipSuccess;
    if ((in_channels != 3 && in_channels != 4) || out_format < 0 || out_format > 5) return hipErrorInvalidValue;
    const unsigned g = nblocks(npix);
    const __half* src = reinterpret_cast<const __half*>(in);
#define RFK_CV(C)                                                                                        \
    switch (out_format) {                                                                               \
        case 0: convert_kernel<C, 0><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
        case 1: convert_kernel<C, 1><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
        case 2: convert_kernel<C, 2><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
        case 3: convert_kernel<C, 3><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
        case 4: convert_kernel<C, 4><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
        default: convert_kernel<C, 5><<<g, kBlock, 0, s>>>(src, out, width, height, flip, swap); break; \
    }
    if (in_channels == 4) {
        RFK_CV(4)
    } else {
        RFK_CV(3)
    }
#undef RFK_CV
    return hipGetLastError();
}

hipError_t density_hdr(const void* in, bool in_half, size_t npix, const unsigned long long* bins97, float* out,
                       hipStream_t s) {
    if (!npix) return hipSuccess;
    if (in_half)
        density_hdr_kernel<true><<<nblocks(npix), kBlock, 0, s>>>(in, npix, bins97, reinterpret_cast<float4*>(out));
    else
        density_hdr_kernel<false><<<nblocks(npix), kBlock, 0, s>>>(in, npix, bins97, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

hipError_t atomic_peak(uint32_t* buf, uint64_t nbins, uint32_t grid, uint32_t block, uint32_t iters, uint32_t seed,
                       hipStream_t s) {
    atomic_peak_kernel<<<grid, block, 0, s>>>(buf, nbins, iters, seed);
    return hipGetLastError();
}

}  // namespace rfk::sk
