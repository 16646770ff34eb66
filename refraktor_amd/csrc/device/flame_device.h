// Device runtime for the iterate kernels: value types, the math overload set
// the variation language targets, the counter-based RNG, the block
// permutation and the fixed-point draw. It is compiled twice: by hiprtc in
// front of every generated flame kernel, and by hipcc into the static kernels
// (bin_points uses the same draw code as the iterate kernel).
//
// Semantics kept from librefrakt's device library
// (assets/kernels/include/refrakt/flamelib.h:4-26, 326-490 and random.h:6-30):
// eps(x) = x == 0 ? 1e-20 : x; sincos -> (sin, cos); cossin -> (cos, sin);
// rand01 = u32 / 4294967295 in the build's precision (in float the divisor
// rounds to 2^32, so the f32 form is the exact product u32 * 2^-32);
// randbit = bit 24; randgauss draws the angle first, then the sum of four
// uniforms.
#pragma once

#ifndef RFK_DEVICE_H
#define RFK_DEVICE_H

typedef unsigned int u32;
typedef unsigned long long u64;
typedef unsigned short u16;
typedef unsigned char u8;

#ifdef RFK_DOUBLE
typedef double Real;
#else
typedef float Real;
#endif

// A sample's parameter row as the iterate kernels read it: constant address
// space (nothing writes it while they run), so loads at workgroup-uniform
// addresses become scalar loads into SGPRs instead of vector loads per pass.
typedef const __attribute__((address_space(4))) Real* param_ptr;

#define RFK_INLINE __device__ __forceinline__

// Start of a block whose arithmetic must come out as written: no FMA
// contraction, no reassociation (needs -ffp-contract=fast-honor-pragmas
// when built with -ffast-math). Used where results are compared bit for bit
// with a host evaluation: xform selection, per-sample sums, binning.
#define RFK_EXACT_FP _Pragma("clang fp contract(off) reassociate(off)")

// correctly rounded products and sums (xform selection)
#ifdef RFK_DOUBLE
#define RFK_MUL_RN(a, b) __dmul_rn((a), (b))
#define RFK_ADD_RN(a, b) __dadd_rn((a), (b))
#else
#define RFK_MUL_RN(a, b) __fmul_rn((a), (b))
#define RFK_ADD_RN(a, b) __fadd_rn((a), (b))
#endif

// ------------------------------------------------------------------ vectors
struct v2 {
    Real x, y;
};
struct v3 {
    Real x, y, z;
};

RFK_INLINE v2 operator+(v2 a, v2 b) { return {a.x + b.x, a.y + b.y}; }
RFK_INLINE v2 operator-(v2 a, v2 b) { return {a.x - b.x, a.y - b.y}; }
RFK_INLINE v2 operator-(v2 a) { return {-a.x, -a.y}; }
RFK_INLINE v2 operator*(v2 a, Real s) { return {a.x * s, a.y * s}; }
RFK_INLINE v2 operator*(Real s, v2 a) { return {s * a.x, s * a.y}; }
RFK_INLINE v2 operator/(v2 a, Real s) { return {a.x / s, a.y / s}; }
RFK_INLINE v2& operator+=(v2& a, v2 b) { a.x += b.x; a.y += b.y; return a; }
RFK_INLINE v2& operator-=(v2& a, v2 b) { a.x -= b.x; a.y -= b.y; return a; }
RFK_INLINE v2& operator*=(v2& a, Real s) { a.x *= s; a.y *= s; return a; }
RFK_INLINE v2& operator/=(v2& a, Real s) { a.x /= s; a.y /= s; return a; }
RFK_INLINE v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RFK_INLINE v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RFK_INLINE v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
RFK_INLINE v3 operator*(v3 a, Real s) { return {a.x * s, a.y * s, a.z * s}; }
RFK_INLINE v3 operator*(Real s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
RFK_INLINE v3 operator/(v3 a, Real s) { return {a.x / s, a.y / s, a.z / s}; }
RFK_INLINE v3& operator+=(v3& a, v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
RFK_INLINE v3& operator-=(v3& a, v3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
RFK_INLINE v3& operator*=(v3& a, Real s) { a.x *= s; a.y *= s; a.z *= s; return a; }
RFK_INLINE v3& operator/=(v3& a, Real s) { a.x /= s; a.y /= s; a.z /= s; return a; }

// ------------------------------------------------------------------- math
namespace fl {

constexpr Real pi = Real(3.14159265358979323846264338327950288);
constexpr Real inv_pi = Real(0.318309886183790671537767526745028724);

#ifdef RFK_DOUBLE
#define RFK_M1(name, f) RFK_INLINE Real name(Real a) { return ::f(a); }
#define RFK_M2(name, f) RFK_INLINE Real name(Real a, Real b) { return ::f(a, b); }
RFK_M1(sin, sin) RFK_M1(cos, cos) RFK_M1(tan, tan) RFK_M1(sinh, sinh) RFK_M1(cosh, cosh) RFK_M1(tanh, tanh)
RFK_M1(asin, asin) RFK_M1(acos, acos) RFK_M1(atan, atan) RFK_M1(asinh, asinh) RFK_M1(acosh, acosh)
RFK_M1(atanh, atanh) RFK_M1(sqrt, sqrt) RFK_M1(rsqrt, rsqrt) RFK_M1(exp, exp) RFK_M1(log, log) RFK_M1(exp2, exp2)
RFK_M1(log2, log2) RFK_M1(exp10, exp10) RFK_M1(log10, log10) RFK_M1(abs, fabs) RFK_M1(floor, floor)
RFK_M1(sinpi, sinpi) RFK_M1(cospi, cospi)
RFK_M2(atan2, atan2) RFK_M2(hypot, hypot) RFK_M2(pow, pow) RFK_M2(modf, fmod) RFK_M2(copysign, copysign)
RFK_INLINE Real tanpi(Real a) { return ::tan(a * pi); }
RFK_INLINE int trunc(Real a) { return int(::trunc(a)); }
RFK_INLINE int rint(Real a) { return int(::rint(a)); }
RFK_INLINE v2 sincos(Real a) { v2 r; ::sincos(a, &r.x, &r.y); return r; }
RFK_INLINE v2 cossin(Real a) { v2 r; ::sincos(a, &r.y, &r.x); return r; }
RFK_INLINE v2 sincospi(Real a) { v2 r; ::sincospi(a, &r.x, &r.y); return r; }
RFK_INLINE v2 cossinpi(Real a) { v2 r; ::sincospi(a, &r.y, &r.x); return r; }
RFK_INLINE Real fma(Real a, Real b, Real c) { return ::fma(a, b, c); }
#define RFK_EPS 1e-20
#else
#define RFK_M1(name, f) RFK_INLINE Real name(Real a) { return ::f(a); }
#define RFK_M2(name, f) RFK_INLINE Real name(Real a, Real b) { return ::f(a, b); }
// The reference compiles flame kernels with fast math
// (src/librefrakt/src/flame_compiler.cpp:861, use_fast_math), under which
// sin, cos, tan, sincos, exp, exp10, log, log2, log10 and pow are the
// hardware approximations; the f32 path does the same with HIP's native
// forms (v_sin/v_cos/v_exp/v_log based, no range reduction), the f64 path
// keeps the library functions.
RFK_M1(sin, __sinf) RFK_M1(cos, __cosf) RFK_M1(tan, __tanf) RFK_M1(sinh, sinhf) RFK_M1(cosh, coshf)
RFK_M1(tanh, tanhf) RFK_M1(asin, asinf) RFK_M1(acos, acosf) RFK_M1(atan, atanf) RFK_M1(asinh, asinhf)
RFK_M1(acosh, acoshf) RFK_M1(atanh, atanhf) RFK_M1(sqrt, sqrtf) RFK_M1(rsqrt, rsqrtf) RFK_M1(exp, __expf)
RFK_M1(log, __logf) RFK_M1(exp2, exp2f) RFK_M1(log2, __log2f) RFK_M1(exp10, __exp10f) RFK_M1(log10, __log10f)
RFK_M1(abs, fabsf) RFK_M1(floor, floorf) RFK_M1(sinpi, sinpif) RFK_M1(cospi, cospif)
RFK_M2(atan2, atan2f) RFK_M2(hypot, hypotf) RFK_M2(pow, __powf) RFK_M2(modf, fmodf) RFK_M2(copysign, copysignf)
RFK_INLINE Real tanpi(Real a) { return ::__tanf(a * pi); }
RFK_INLINE int trunc(Real a) { return int(::truncf(a)); }
RFK_INLINE int rint(Real a) { return int(::rintf(a)); }
RFK_INLINE v2 sincos(Real a) { v2 r; ::__sincosf(a, &r.x, &r.y); return r; }
RFK_INLINE v2 cossin(Real a) { v2 r; ::__sincosf(a, &r.y, &r.x); return r; }
RFK_INLINE v2 sincospi(Real a) { v2 r; ::sincospif(a, &r.x, &r.y); return r; }
RFK_INLINE v2 cossinpi(Real a) { v2 r; ::sincospif(a, &r.y, &r.x); return r; }
RFK_INLINE Real fma(Real a, Real b, Real c) { return ::fmaf(a, b, c); }
#define RFK_EPS 1e-20f
#endif

// elementwise vector forms of the one-argument functions
#define RFK_VEC1(name)                                                           \
    RFK_INLINE v2 name(v2 a) { return {name(a.x), name(a.y)}; }                   \
    RFK_INLINE v3 name(v3 a) { return {name(a.x), name(a.y), name(a.z)}; }
RFK_VEC1(sin) RFK_VEC1(cos) RFK_VEC1(tan) RFK_VEC1(sinh) RFK_VEC1(cosh) RFK_VEC1(sqrt) RFK_VEC1(rsqrt)
RFK_VEC1(exp) RFK_VEC1(log) RFK_VEC1(log2) RFK_VEC1(log10) RFK_VEC1(abs) RFK_VEC1(acos) RFK_VEC1(sinpi)
RFK_VEC1(cospi)

RFK_INLINE Real eps(Real a) { return (a == Real(0)) ? Real(RFK_EPS) : a; }
RFK_INLINE v2 eps(v2 a) { return {eps(a.x), eps(a.y)}; }
RFK_INLINE v3 eps(v3 a) { return {eps(a.x), eps(a.y), eps(a.z)}; }

RFK_INLINE v2 fma(Real a, v2 b, v2 c) { return {fma(a, b.x, c.x), fma(a, b.y, c.y)}; }
RFK_INLINE v2 fma(v2 a, Real b, v2 c) { return {fma(a.x, b, c.x), fma(a.y, b, c.y)}; }
RFK_INLINE v3 fma(Real a, v3 b, v3 c) { return {fma(a, b.x, c.x), fma(a, b.y, c.y), fma(a, b.z, c.z)}; }
RFK_INLINE v3 fma(v3 a, Real b, v3 c) { return {fma(a.x, b, c.x), fma(a.y, b, c.y), fma(a.z, b, c.z)}; }

RFK_INLINE bool is_even(int x) { return !(x & 1); }
RFK_INLINE v2 elmul(v2 a, v2 b) { return {a.x * b.x, a.y * b.y}; }
RFK_INLINE v2 vec2(Real x, Real y) { return {x, y}; }
RFK_INLINE v3 vec3(Real x, Real y, Real z) { return {x, y, z}; }

// affine in librefrakt field order (a, d, b, e, c, f): x' = a x + b y + c,
// y' = d x + e y + f, as two nested FMAs (assets/templates/flame.tpl:10-14)
template <class Ptr>
RFK_INLINE v2 affine(Ptr A, v2 p) {
    return {fma(A[0], p.x, fma(A[2], p.y, A[4])), fma(A[1], p.x, fma(A[3], p.y, A[5]))};
}

}  // namespace fl

// ---------------------------------------------------------- Philox4x32-10
// Salmon et al., "Parallel random numbers: as easy as 1, 2, 3" (SC'11):
// multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 /
// 0xBB67AE85, ten rounds.
struct u32x4 {
    u32 x, y, z, w;
};

RFK_INLINE u32x4 philox4x32_10(u32x4 c, u32 k0, u32 k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        u32 lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        u32 lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = {hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    }
    return c;
}

// Stream tags: the fourth counter word separates the independent streams.
#define RFK_TAG_PICK 1u    // per (pass/4, block, wave): wave xform choice
#define RFK_TAG_PERM 2u    // per (pass/2, block): block permutation keys
#define RFK_TAG_LANE 3u    // per (pass, walker thread, k): variation draws
#define RFK_TAG_INIT 4u    // per walker thread: initial colour
#define RFK_TAG_JITTER 5u  // per (sample, k): anti-alias jitter

#ifdef RFK_DOUBLE
RFK_INLINE Real rfk_u01(u32 u) { return double(u) / 4294967295.0; }
#else
RFK_INLINE Real rfk_u01(u32 u) { return float(u) * 2.3283064365386963e-10f; }
#endif

// Per-walker draw stream for one pass: word j of the pass is word j%4 of
// philox({pass, gid, j/4, tag}, key), generated lazily four at a time.
struct rng_t {
    u32 k0, k1, c0, c1, c2, tag;
    u32 b0, b1, b2, b3, avail;

    RFK_INLINE void init(u32 key0, u32 key1, u32 counter1, u32 stream_tag) {
        k0 = key0; k1 = key1; c0 = 0; c1 = counter1; c2 = 0; tag = stream_tag; avail = 0;
        b0 = b1 = b2 = b3 = 0;
    }
    RFK_INLINE void begin(u32 counter0) { c0 = counter0; c2 = 0; avail = 0; }
    RFK_INLINE u32 next() {
        if (avail == 0) {
            u32x4 r = philox4x32_10({c0, c1, c2, tag}, k0, k1);
            c2++;
            b0 = r.x; b1 = r.y; b2 = r.z; b3 = r.w;
            avail = 4;
        }
        u32 v = b0;
        b0 = b1; b1 = b2; b2 = b3;
        avail--;
        return v;
    }
    RFK_INLINE Real rand01() { return rfk_u01(next()); }
    RFK_INLINE Real randbit() { return Real((next() >> 24) & 1u); }
    RFK_INLINE v2 randgauss(Real sd = Real(1)) {
        v2 sc = fl::sincospi(rand01() * Real(2));
        Real a = rand01(), b = rand01(), c = rand01(), d = rand01();
        Real r = sd * (((a + b) + c) + d - Real(2));
        return {r * sc.y, r * sc.x};
    }
};

// ------------------------------------------------------- block permutation
// A bijection on [0, 2^bits): two rounds of (multiply by an odd constant, add)
// mod 2^bits followed by an xorshift, keyed by two 32-bit words.
RFK_INLINE u32 rfk_permute(u32 t, u32 wa, u32 wb, u32 bits) {
    const u32 mask = (1u << bits) - 1u;
    const u32 sh = bits > 1u ? bits >> 1 : 1u;
    u32 x = (t * ((wa & 0xFFFFu) | 1u) + (wa >> 16)) & mask;
    x ^= x >> sh;
    x = (x * ((wb & 0xFFFFu) | 1u) + (wb >> 16)) & mask;
    x ^= x >> sh;
    return x;
}

// --------------------------------------------------------------- badvalue
// flamelib.h:13,25 badvalue: NaN, +-inf or |x| > 1e10. Tested on the bit
// pattern so fast-math cannot fold the NaN case away.
RFK_INLINE bool rfk_badvalue(float x) { return (__float_as_uint(x) & 0x7FFFFFFFu) > 0x501502F9u; }
#ifdef RFK_DOUBLE
RFK_INLINE bool rfk_badvalue(double x) {
    return (__double_as_longlong(x) & 0x7FFFFFFFFFFFFFFFll) > 0x4202A05F20000000ll;
}
#endif

// ------------------------------------------------------------- hammersley
// refactor.cu:10-55: (i/N, bitrev(i)/N) mapped to [-1, 1]^2, N rounded up to
// a power of two when odd.
RFK_INLINE u32 rfk_bitrev(u32 n) { return __brev(n); }
RFK_INLINE v2 rfk_hammersley(u32 idx, u32 total) {
    u32 max_val = total;
    if (total % 2u != 0u) {
        u32 v = total - 1u;
        v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
        max_val = v + 1u;
    }
    u32 radix = 0, v = max_val;
    while (v >>= 1) radix++;
    Real inv_max = Real(1.0) / Real(max_val);
    u32 flipped = radix ? (rfk_bitrev(idx) >> (32u - radix)) : 0u;
    return {Real(idx) * inv_max * Real(2.0) - Real(1.0), Real(flipped) * inv_max * Real(2.0) - Real(1.0)};
}

// ------------------------------------------------------------------- draw
// The fixed-point bin update (replaces the racy float4 read-modify-write of
// refactor.cu:347-386). One draw adds, to the bin (x, y) row-major:
//   R,G,B += round(lerp(palette[floor(i)], palette[ceil(i)], frac(i)) * opacity)
//   W     += round(255 * opacity)                     with i = 255 * clamp(z, 0, 1)
// so bin / 255 is the reference's float4 bin in draw units. All roundings are
// spelled out (explicit fmaf, __fmul_rn) so any compiler produces the same bits.
struct rfk_draw_t {
    u32 pixel, px, py;
    u32 r, g, b, w;
    bool hit;
};

RFK_INLINE rfk_draw_t rfk_make_draw(const float* __restrict__ S, const u32* pal, u32 width, u32 height, float x,
                                    float y, float z, float opacity) {
    RFK_EXACT_FP
    rfk_draw_t d;
    d.hit = false;
    float sx = truncf(fmaf(S[0], x, fmaf(S[2], y, S[4])));
    float sy = truncf(fmaf(S[1], x, fmaf(S[3], y, S[5])));
    if (sx >= 0.0f && sy >= 0.0f && sx < float(width) && sy < float(height) && opacity > 0.0f) {
        d.hit = true;
        d.px = u32(sx);
        d.py = u32(sy);
        d.pixel = d.py * width + d.px;
        float zc = fminf(fmaxf(z, 0.0f), 1.0f);
        float idx = __fmul_rn(zc, 255.0f);
        float mix = idx - truncf(idx);
        u32 lo = pal[u32(floorf(idx))], hi = pal[u32(ceilf(idx))];
        float lr = float(lo & 0xFFu), lg = float((lo >> 8) & 0xFFu), lb = float((lo >> 16) & 0xFFu);
        float hr = float(hi & 0xFFu), hg = float((hi >> 8) & 0xFFu), hb = float((hi >> 16) & 0xFFu);
        float r = fmaf(mix, hr - lr, lr), g = fmaf(mix, hg - lg, lg), b = fmaf(mix, hb - lb, lb);
        d.r = u32(fmaf(r, opacity, 0.5f));
        d.g = u32(fmaf(g, opacity, 0.5f));
        d.b = u32(fmaf(b, opacity, 0.5f));
        d.w = u32(fmaf(opacity, 255.0f, 0.5f));
    }
    return d;
}

// The palette as the iterate kernels' draws read it (in LDS): entry i's
// channels as floats and their step to entry i + 1 (0 for the last entry).
// The small integers and their differences are exact in float, and at a
// whole index (mix 0) fmaf(0, step, lo) is lo whatever the step, so the draw
// below gives rfk_make_draw's bits with one entry read and no unpacking.
struct alignas(16) rfk_pal_lut {
    float lo[4], step[4];
};

RFK_INLINE rfk_pal_lut rfk_pal_lut_entry(const u32* pal, u32 i) {
    const u32 lo = pal[i], hi = pal[i < 255u ? i + 1u : i];
    rfk_pal_lut e;
#pragma unroll
    for (u32 c = 0; c < 3u; c++) {
        const float l = float((lo >> (8u * c)) & 0xFFu), h = float((hi >> (8u * c)) & 0xFFu);
        e.lo[c] = l;
        e.step[c] = h - l;
    }
    e.lo[3] = 0.0f;
    e.step[3] = 0.0f;
    return e;
}

RFK_INLINE rfk_draw_t rfk_make_draw(const float* __restrict__ S, const rfk_pal_lut* lut, u32 width, u32 height,
                                    float x, float y, float z, float opacity) {
    RFK_EXACT_FP
    rfk_draw_t d;
    d.hit = false;
    float sx = truncf(fmaf(S[0], x, fmaf(S[2], y, S[4])));
    float sy = truncf(fmaf(S[1], x, fmaf(S[3], y, S[5])));
    if (sx >= 0.0f && sy >= 0.0f && sx < float(width) && sy < float(height) && opacity > 0.0f) {
        d.hit = true;
        d.px = u32(sx);
        d.py = u32(sy);
        d.pixel = d.py * width + d.px;
        float zc = fminf(fmaxf(z, 0.0f), 1.0f);
        float idx = __fmul_rn(zc, 255.0f);
        float fi = floorf(idx);  // idx >= 0: floor = trunc
        float mix = idx - fi;
        const rfk_pal_lut e = lut[u32(fi)];
        float r = fmaf(mix, e.step[0], e.lo[0]), g = fmaf(mix, e.step[1], e.lo[1]), b = fmaf(mix, e.step[2], e.lo[2]);
        d.r = u32(fmaf(r, opacity, 0.5f));
        d.g = u32(fmaf(g, opacity, 0.5f));
        d.b = u32(fmaf(b, opacity, 0.5f));
        d.w = u32(fmaf(opacity, 255.0f, 0.5f));
    }
    return d;
}

// 64-bit bins: four u64 atomics, no wrap at any budget
RFK_INLINE void rfk_commit_draw(u64* __restrict__ hist, const rfk_draw_t& d) {
    u64* h = hist + 4ull * d.pixel;
    atomicAdd(h + 0, u64(d.r));
    atomicAdd(h + 1, u64(d.g));
    atomicAdd(h + 2, u64(d.b));
    atomicAdd(h + 3, u64(d.w));
}

// 32-bit bins: four u32 atomics, or two u64 atomics on (R,G) and (B,W) pairs
// (a channel carries into its neighbour only past 2^32, where u32 bins have
// wrapped anyway)
RFK_INLINE void rfk_commit_draw(u32* __restrict__ hist, const rfk_draw_t& d) {
#if defined(RFK_ATOMIC64) && RFK_ATOMIC64
    u64* h64 = reinterpret_cast<u64*>(hist) + 2ull * d.pixel;
    atomicAdd(h64, u64(d.r) | (u64(d.g) << 32));
    atomicAdd(h64 + 1, u64(d.b) | (u64(d.w) << 32));
#else
    u32* h = hist + 4ull * d.pixel;
    atomicAdd(h + 0, d.r);
    atomicAdd(h + 1, d.g);
    atomicAdd(h + 2, d.b);
    atomicAdd(h + 3, d.w);
#endif
}

// --------------------------------------------------------- kernel ABI
// Launch arguments of the generated kernels, passed by value. The host side
// includes this same definition.
struct rfk_iter_args {
    const float* params;       // [samples][param_stride] flame parameters per block
    const u32* palettes;       // [samples][256] RGBA8
    float* wx;                 // [walkers]
    float* wy;
    float* wc;
    u32* block_pass;           // [blocks] passes completed by each block
    u32* hist;                 // [height][width][4] u32, or u64 when the kernel has RFK_HIST64
    const volatile u32* stop_flag;
    u64* counters;             // [8] hits, passes, draws, bad, min start, max stop, stopped blocks, -
    float* trace;              // [npasses][walkers][4] or null
    u32* trace_xid;            // [npasses][walkers/64] or null
    u32* wlast;                // [walkers] last xform per walker (xaos) or null
    const u32* row_stream;     // precalc: stream id of each params row (null: a.stream)
    const u32* row_local;      // precalc: jitter counter of each row (null: the row index)
    u64 time_limit_ticks;      // 0 = none
    u64 quality_target;        // per block, in 255*opacity units; 0 = none
    u32 width, height;
    u32 param_stride;
    u32 seed, stream;
    u32 npasses;               // passes to run in this launch (per walker)
    u32 total_walkers;
    u32 flags;                 // RFK_FLAG_*
    u32 check_every;           // passes between bailout checks (power of two)
    u32 nsamples;              // rows of params (precalc kernel)
    u64* queue;                // queue binning: [blocks][npasses][RFK_TPB] draw records
    u32* qcounts;              // queue binning: [tiles][blocks] records per screen tile
    u32 tiles_x;               // queue binning: 64-pixel tiles per row
    u32 ntiles;                // queue binning: tiles in the image
};

// Persistent batch launch: work items are (frame, block) pairs
struct rfk_batch_args {
    u32* queue;                 // work-item counter (zeroed before the launch)
    const u32* item_frame;      // [nitems]
    const u32* item_block;      // [nitems] block inside its frame
    const u32* frame_row0;      // [frames] first params/palette row of the frame
    const u32* frame_blocks;    // [frames]
    const u32* frame_passes;    // [frames] binned passes per walker
    u64* frame_counters;        // [frames][4] hits, draws, passes, bad
    u32 nitems;
    u32 fuse;                   // un-binned passes after seeding
    u32 frame_pixels;           // width * height
    u32 pad;
};

// Queue-binned batch chunk: slots [slot0, slot0 + grid / cols) of `slots`
struct rfk_qbatch_args {
    const u32* slots;           // [slots][4] frame, first pass, binned passes, fuse passes
    u64* frame_counters;        // [frames][4] hits, draws, passes, bad
    u32 slot0;                  // first slot of this chunk
    u32 cols;                   // walker blocks per frame
    u32 passes;                 // record passes per block in this chunk
    u32 frame_pixels;           // width * height
};

#define RFK_FLAG_DRAW 1u
#define RFK_FLAG_INIT 2u
#define RFK_FLAG_TRACE 4u
#define RFK_FLAG_BAILOUT 8u

#define RFK_CNT_HITS 0
#define RFK_CNT_PASSES 1
#define RFK_CNT_DRAWS 2
#define RFK_CNT_BAD 3
#define RFK_CNT_START 4
#define RFK_CNT_STOP 5
#define RFK_CNT_STOPPED 6
#define RFK_CNT_ORIGIN 7   // queue binning: device clock at the first block of a bin call

#endif  // RFK_DEVICE_H
