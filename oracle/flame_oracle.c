/*
 * ORACLE — test infrastructure only. Nothing in the product imports, links
 * or calls this file; tests/, the smoke check and bench.py's cpu_baseline use
 * it as the checker and as the timed host-core baseline.
 *
 * A scalar CPU restatement of the chaos-game iterate/bin path:
 *   - Philox4x32-10 (Salmon et al., SC'11) and the stream layout of the GPU
 *     path (tags PICK/PERM/LANE/INIT);
 *   - xform selection by cumulative weight (assets/templates/flame.tpl:110-123),
 *     float arithmetic, so choices are bit-identical;
 *   - the variations used by the test and benchmark flames, hand-written in
 *     real from the standard flam3 formulas (not from the product's DSL);
 *   - the per-pass loop of the reference (refactor.cu:176-210, 347-386):
 *     xform, bad-value reset, final xform on a copy, draw, then the block
 *     permutation of the GPU scheme;
 *   - the fixed-point RGBA draw, with the same explicit roundings as the
 *     GPU (compiled with -ffp-contract=off), so point-list histograms are
 *     bit-identical;
 *   - an OpenMP render used as the host-core baseline.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* The iterate path's float type: double for the checker. The host-core
 * baseline build (oracle/build.py: -DORC_REAL=float -O3 -march=native
 * -fsingle-precision-constant) runs the same loop in single precision with
 * the type-generic functions of <tgmath.h>. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
#include <tgmath.h>

typedef uint32_t u32;
typedef uint64_t u64;

/* ------------------------------------------------------------ Philox */
void orc_philox(const u32 ctr[4], const u32 key[2], u32 out[4]) {
    u32 c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    u32 k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        u64 p0 = (u64)0xD2511F53u * c0, p1 = (u64)0xCD9E8D57u * c2;
        u32 hi0 = (u32)(p0 >> 32), lo0 = (u32)p0, hi1 = (u32)(p1 >> 32), lo1 = (u32)p1;
        u32 n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define TAG_PICK 1u
#define TAG_PERM 2u
#define TAG_LANE 3u
#define TAG_INIT 4u

static inline float u01(u32 u) { return (float)u * 2.3283064365386963e-10f; }

typedef struct {
    u32 k0, k1, c0, c1, c2, tag;
    u32 buf[4];
    int avail, pos;
    const u32* list; /* when set: the words are list[0..nlist) in order */
    int nlist, lpos;
} lane_rng;

static inline void rng_init(lane_rng* r, u32 k0, u32 k1, u32 c1, u32 tag) {
    r->k0 = k0; r->k1 = k1; r->c0 = 0; r->c1 = c1; r->c2 = 0; r->tag = tag; r->avail = 0; r->pos = 0;
    r->list = NULL; r->nlist = 0; r->lpos = 0;
}
static inline void rng_begin(lane_rng* r, u32 pass) { r->c0 = pass; r->c2 = 0; r->avail = 0; r->pos = 0; }
static inline u32 rng_next(lane_rng* r) {
    if (r->list) return r->lpos < r->nlist ? r->list[r->lpos++] : 0u;
    if (r->avail == 0) {
        u32 ctr[4] = {r->c0, r->c1, r->c2, r->tag}, key[2] = {r->k0, r->k1};
        orc_philox(ctr, key, r->buf);
        r->c2++;
        r->avail = 4;
        r->pos = 0;
    }
    r->avail--;
    return r->buf[r->pos++];
}
static inline real rand01(lane_rng* r) { return (real)u01(rng_next(r)); }
static inline real randbit(lane_rng* r) { return (real)((rng_next(r) >> 24) & 1u); }
static void randgauss(lane_rng* r, real sd, real* gx, real* gy) {
    real a = (real)(u01(rng_next(r)) * 2.0f);
    real s = sin(M_PI * a), c = cos(M_PI * a);
    float f0 = u01(rng_next(r)), f1 = u01(rng_next(r)), f2 = u01(rng_next(r)), f3 = u01(rng_next(r));
    real rr = sd * ((real)f0 + f1 + f2 + f3 - 2.0);
    *gx = rr * c;
    *gy = rr * s;
}

/* ------------------------------------------------------- permutation */
static inline u32 permute(u32 t, u32 wa, u32 wb, u32 bits) {
    const u32 mask = (1u << bits) - 1u;
    const u32 sh = bits > 1u ? bits >> 1 : 1u;
    u32 x = (t * ((wa & 0xFFFFu) | 1u) + (wa >> 16)) & mask;
    x ^= x >> sh;
    x = (x * ((wb & 0xFFFFu) | 1u) + (wb >> 16)) & mask;
    x ^= x >> sh;
    return x;
}

/* ------------------------------------------------------- variations */
enum {
    V_LINEAR = 0, V_SINUSOIDAL, V_SPHERICAL, V_SWIRL, V_HORSESHOE, V_POLAR, V_HANDKERCHIEF, V_HEART, V_DISC,
    V_SPIRAL, V_HYPERBOLIC, V_DIAMOND, V_EX, V_JULIA, V_BUBBLE, V_CURL, V_FISHEYE, V_EYEFISH, V_BLUR,
    V_GAUSSIAN_BLUR, V_WAVES, V_JULIAN, V_BENT, V_CYLINDER, V_EXPONENTIAL, V_POWER, V_COSINE, V_NOISE,
    V_RECTANGLES, V_PDJ, V_CROSS, V_EXP, V_LOG, V_SINH, V_COSH, V_SIN, V_COS, V_TANGENT, V_SQUARE, V_POLAR2,
    V_FOCI, V_PERSPECTIVE, V_BLOB, V_RINGS, V_FAN, V_BENT2, V_NGON, V_COUNT
};

/* Program layout (ints), per xform (standard xforms, then the final one):
 *   n_vlinks, off_weight, off_color, off_speed, off_opacity,
 *   per vlink: off_affine, n_vars, per var: var_id, off_weight, n_params, param offsets...
 * Parameter offsets are listed in sorted order of ALL the variation's
 * parameter names (user and precalc together); apply_variation indexes prm[]
 * in that order. */

typedef struct {
    const int* prog;
    int n_xforms, has_final;
    int chaos_base;  /* offset of the xaos rows (n x (1 + n) reals), -1 without xaos */
    int xform_start[256];
} flame_prog;

static void prog_index(flame_prog* fp, const int* prog, int n_xforms, int has_final, int chaos_base) {
    fp->prog = prog;
    fp->n_xforms = n_xforms;
    fp->has_final = has_final;
    fp->chaos_base = chaos_base;
    int i = 0;
    for (int x = 0; x < n_xforms + has_final; x++) {
        fp->xform_start[x] = i;
        int nv = prog[i];
        i += 5;
        for (int v = 0; v < nv; v++) {
            int nvar = prog[i + 1];
            i += 2;
            for (int k = 0; k < nvar; k++) {
                int np = prog[i + 2];
                i += 3 + np;
            }
        }
    }
}

static inline real eps_(real x) { return x == 0.0 ? 1e-20 : x; }

static void apply_variation(int id, real w, const real* prm, const real* aff, real px, real py,
                            real* rx, real* ry, lane_rng* rs) {
    real sumsq = px * px + py * py, r = hypot(px, py);
    real atn = atan2(px, py), phi = atan2(py, px);
    real rinv = 1.0 / sqrt(sumsq);
    real sina = px * rinv, cosa = py * rinv;
    switch (id) {
        case V_LINEAR: *rx += w * px; *ry += w * py; break;
        case V_SINUSOIDAL: *rx += w * sin(px); *ry += w * sin(py); break;
        case V_SPHERICAL: { real k = w / eps_(sumsq); *rx += k * px; *ry += k * py; } break;
        case V_SWIRL: {
            real s = sin(sumsq), c = cos(sumsq);
            *rx += w * (s * px - c * py);
            *ry += w * (c * px + s * py);
        } break;
        case V_HORSESHOE: { real k = w / eps_(r); *rx += k * (px - py) * (px + py); *ry += k * 2.0 * px * py; } break;
        case V_POLAR: *rx += w * atn / M_PI; *ry += w * (r - 1.0); break;
        case V_HANDKERCHIEF: *rx += w * r * sin(atn + r); *ry += w * r * cos(atn - r); break;
        case V_HEART: { real a = r * atn; *rx += w * r * sin(a); *ry -= w * r * cos(a); } break;
        case V_DISC: { real k = w * atn / M_PI; *rx += k * sin(M_PI * r); *ry += k * cos(M_PI * r); } break;
        case V_SPIRAL: {
            real re = eps_(r);
            *rx += w / re * (cosa + sin(re));
            *ry += w / re * (sina - cos(re));
        } break;
        case V_HYPERBOLIC: { real re = eps_(r); *rx += w * sina / re; *ry += w * cosa * re; } break;
        case V_DIAMOND: *rx += w * sina * cos(r); *ry += w * cosa * sin(r); break;
        case V_EX: {
            real n0 = sin(atn + r), n1 = cos(atn - r);
            real m0 = n0 * n0 * n0 * r, m1 = n1 * n1 * n1 * r;
            *rx += w * (m0 + m1);
            *ry += w * (m0 - m1);
        } break;
        case V_JULIA: {
            real a = 0.5 * atn + M_PI * randbit(rs);
            real k = w * sqrt(r);
            *rx += k * cos(a);
            *ry += k * sin(a);
        } break;
        case V_BUBBLE: { real k = w / (0.25 * sumsq + 1.0); *rx += k * px; *ry += k * py; } break;
        case V_CURL: {
            real c1 = prm[0], c2 = prm[1];
            real re = 1.0 + c1 * px + c2 * (px * px - py * py);
            real im = c1 * py + 2.0 * c2 * px * py;
            real k = w / (re * re + im * im);
            *rx += (px * re + py * im) * k;
            *ry += (py * re - px * im) * k;
        } break;
        case V_FISHEYE: { real k = 2.0 * w / (r + 1.0); *rx += k * py; *ry += k * px; } break;
        case V_EYEFISH: { real k = 2.0 * w / (r + 1.0); *rx += k * px; *ry += k * py; } break;
        case V_BLUR: {
            real k = w * rand01(rs);
            real a = (real)((float)rand01(rs) * 2.0f);
            *rx += k * cos(M_PI * a);
            *ry += k * sin(M_PI * a);
        } break;
        case V_GAUSSIAN_BLUR: { real gx, gy; randgauss(rs, 1.0, &gx, &gy); *rx += w * gx; *ry += w * gy; } break;
        case V_WAVES: /* params: dx2, dy2 (precalc) */
            *rx += w * (px + aff[2] * sin(py * prm[0]));
            *ry += w * (py + aff[3] * sin(px * prm[1]));
            break;
        case V_JULIAN: { /* params: cn, dist, power, rN */
            real cn = prm[0], power = prm[2], rN = prm[3];
            real k = trunc(rN * rand01(rs));
            real a = (phi + 2.0 * M_PI * k) / power;
            real m = w * pow(sumsq, cn);
            *rx += m * cos(a);
            *ry += m * sin(a);
        } break;
        case V_BENT: {
            real qx = px < 0.0 ? px * 2.0 : px, qy = py < 0.0 ? py * 0.5 : py;
            *rx += w * qx; *ry += w * qy;
        } break;
        case V_CYLINDER: *rx += w * sin(px); *ry += w * py; break;
        case V_EXPONENTIAL: { real e = w * exp(px - 1.0); *rx += e * cos(M_PI * py); *ry += e * sin(M_PI * py); } break;
        case V_POWER: { real k = w * pow(r, sina); *rx += k * cosa; *ry += k * sina; } break;
        case V_COSINE: *rx += w * cos(M_PI * px) * cosh(py); *ry -= w * sin(M_PI * px) * sinh(py); break;
        case V_NOISE: {
            real k = w * rand01(rs);
            real a = (real)(2.0f * (float)rand01(rs));
            *rx += k * px * cos(M_PI * a);
            *ry += k * py * sin(M_PI * a);
        } break;
        case V_RECTANGLES: { /* x, y */
            real ax = prm[0], ay = prm[1];
            *rx += ax == 0.0 ? w * px : w * ((2.0 * floor(px / ax) + 1.0) * ax - px);
            *ry += ay == 0.0 ? w * py : w * ((2.0 * floor(py / ay) + 1.0) * ay - py);
        } break;
        case V_PDJ: /* a b c d */
            *rx += w * (sin(prm[0] * py) - cos(prm[1] * px));
            *ry += w * (sin(prm[2] * px) - cos(prm[3] * py));
            break;
        case V_CROSS: { real s = px * px - py * py; real k = w / sqrt(eps_(s * s)); *rx += k * px; *ry += k * py; } break;
        case V_EXP: { real e = w * exp(px); *rx += e * cos(py); *ry += e * sin(py); } break;
        case V_LOG: *rx += 0.5 * w * log(sumsq); *ry += w * phi; break;
        /* sinh and cos as librefrakt's programs (config/variations_new.yml:880-885,
         * 841-846): sinh takes the hyperbolic factors of y, cos adds the
         * imaginary part; flam3 publishes sinh(x), cosh(x) and a minus */
        case V_SINH: *rx += w * sinh(py) * cos(py); *ry += w * cosh(py) * sin(py); break;
        case V_COSH: *rx += w * cosh(px) * cos(py); *ry += w * sinh(px) * sin(py); break;
        case V_SIN: *rx += w * sin(px) * cosh(py); *ry += w * cos(px) * sinh(py); break;
        case V_COS: *rx += w * cos(px) * cosh(py); *ry += w * sin(px) * sinh(py); break;
        case V_TANGENT: *rx += w * sin(px) / cos(py); *ry += w * tan(py); break;
        case V_SQUARE: { real a = rand01(rs), b = rand01(rs); *rx += w * (a - 0.5); *ry += w * (b - 0.5); } break;
        case V_POLAR2: *rx += w / M_PI * atn; *ry += 0.5 * w / M_PI * log(sumsq); break;
        case V_FOCI: {
            real ex = 0.5 * exp(px), enx = 0.25 / ex;
            real k = w / (ex + enx - cos(py));
            *rx += k * (ex - enx);
            *ry += k * sin(py);
        } break;
        case V_PERSPECTIVE: { /* angle dist vfcos vsin */
            real dist = prm[1], vfcos = prm[2], vsin = prm[3];
            real t = 1.0 / (dist - py * vsin);
            *rx += w * dist * px * t;
            *ry += w * vfcos * py * t;
        } break;
        case V_BLOB: { /* high low waves */
            real hi = prm[0], lo = prm[1], wv = prm[2];
            real k = w * r * (lo + (hi - lo) * (0.5 + 0.5 * sin(wv * atn)));
            *rx += k * sina;
            *ry += k * cosa;
        } break;
        case V_RINGS: {
            real d = eps_(aff[4] * aff[4]);
            real k = fmod(r + d, 2.0 * d) - d + r * (1.0 - d);
            *rx += w * k * cosa;
            *ry += w * k * sina;
        } break;
        case V_FAN: {
            real d = M_PI * eps_(aff[4] * aff[4]), h = 0.5 * d;
            real a = atn + copysign(h, h - fmod(atn + aff[5], d));
            *rx += w * r * cos(a);
            *ry += w * r * sin(a);
        } break;
        case V_BENT2: {
            real qx = px < 0.0 ? px * prm[0] : px, qy = py < 0.0 ? py * prm[1] : py;
            *rx += w * qx; *ry += w * qy;
        } break;
        case V_NGON: { /* circle corners power sides */
            real circle = prm[0], corners = prm[1], power = prm[2], sides = prm[3];
            real rf = pow(sumsq, 0.5 * power);
            real b = 2.0 * M_PI / sides;
            real ph = phi - b * floor(phi / b);
            if (ph > 0.5 * b) ph -= b;
            real amp = (corners * (1.0 / eps_(cos(ph)) - 1.0) + circle) / eps_(rf);
            *rx += w * px * amp;
            *ry += w * py * amp;
        } break;
        default: break;
    }
}

/* One variation at n points p (after the vlink affine): prm in the sorted
 * order of all the variation's parameter names, aff as (a, d, b, e, c, f),
 * words: nwords random words per point, consumed in order. */
void orc_eval_variation(int id, real w, const real* prm, const real* aff, size_t n, const real* px,
                        const real* py, const u32* words, int nwords, real* rx, real* ry) {
    for (size_t i = 0; i < n; i++) {
        lane_rng rs;
        rng_init(&rs, 0u, 0u, 0u, 0u);
        rs.list = words + i * (size_t)nwords;
        rs.nlist = nwords;
        real x = 0.0, y = 0.0;
        apply_variation(id, w, prm, aff, px[i], py[i], &x, &y, &rs);
        rx[i] = x;
        ry[i] = y;
    }
}

/* affine in (a, d, b, e, c, f) order; aff[] below is passed as (a,d,b,e,c,f) */
static real apply_xform(const flame_prog* fp, int x, const float* P, real* px, real* py, real* pz,
                          lane_rng* rs) {
    const int* q = fp->prog + fp->xform_start[x];
    int nv = q[0];
    real w_color = P[q[2]], w_speed = P[q[3]], opacity = P[q[4]];
    q += 5;
    real ix = *px, iy = *py;
    for (int v = 0; v < nv; v++) {
        int aoff = q[0], nvar = q[1];
        q += 2;
        real aff[6];
        for (int k = 0; k < 6; k++) aff[k] = P[aoff + k];
        /* x' = a x + b y + c ; y' = d x + e y + f */
        real ax = aff[0] * ix + aff[2] * iy + aff[4];
        real ay = aff[1] * ix + aff[3] * iy + aff[5];
        real rx = 0.0, ry = 0.0;
        real affv[6] = {aff[0], aff[1], aff[2], aff[3], aff[4], aff[5]};
        for (int k = 0; k < nvar; k++) {
            int id = q[0], woff = q[1], np = q[2];
            real prm[16];
            for (int j = 0; j < np && j < 16; j++) prm[j] = P[q[3 + j]];
            /* waves/popcorn-style access to the affine uses b and e at [2],[3]
             * through the vlink affine (a,d,b,e,c,f) -> pass b,e and c,f */
            real affx[6] = {affv[0], affv[1], affv[2], affv[3], affv[4], affv[5]};
            apply_variation(id, P[woff], prm, affx, ax, ay, &rx, &ry, rs);
            q += 3 + np;
        }
        ix = rx;
        iy = ry;
    }
    *px = ix;
    *py = iy;
    *pz = *pz * (1.0 - w_speed) + w_color * w_speed;
    return opacity;
}

static u32 select_xform(const flame_prog* fp, const float* P, float ratio) {
    const int* q;
    float rsum = 0.0f;
    u32 last_nonzero = 0;
    ratio = ratio * P[12];
    for (int i = 0; i < fp->n_xforms; i++) {
        q = fp->prog + fp->xform_start[i];
        float w = P[q[1]];
        if (i + 1 < fp->n_xforms) {
            if (w != 0.0f && rsum + w >= ratio) return (u32)i;
            rsum = rsum + w;
            if (w != 0.0f) last_nonzero = (u32)i;
        } else {
            return w != 0.0f ? (u32)i : last_nonzero;
        }
    }
    return last_nonzero;
}

/* xaos selection (flame.tpl:127-142): row `last` of the xaos table (its
 * first real is the row's weight sum), cumulative scan as select_xform;
 * no previous xform yet (last >= n) uses the plain weights */
static u32 select_xform_chaos(const flame_prog* fp, const float* P, u32 last, float ratio) {
    const int n = fp->n_xforms;
    if (last >= (u32)n) return select_xform(fp, P, ratio);
    const float* row = P + fp->chaos_base + (size_t)last * (n + 1);
    float rsum = 0.0f;
    u32 last_nonzero = 0;
    ratio = ratio * row[0];
    for (int i = 0; i < n; i++) {
        float w = row[i + 1];
        if (i + 1 < n) {
            if (w != 0.0f && rsum + w >= ratio) return (u32)i;
            rsum = rsum + w;
            if (w != 0.0f) last_nonzero = (u32)i;
        } else {
            return w != 0.0f ? (u32)i : last_nonzero;
        }
    }
    return last_nonzero;
}

static inline int badvalue(real v) { return !(fabs(v) <= 1e10); }

/* -------------------------------------------------------------- draw */
static inline int make_draw(const float* S, const u32* pal, u32 width, u32 height, float x, float y, float z,
                            float opacity, u32* pixel, u32 inc[4]) {
    float sx = truncf(fmaf(S[0], x, fmaf(S[2], y, S[4])));
    float sy = truncf(fmaf(S[1], x, fmaf(S[3], y, S[5])));
    if (!(sx >= 0.0f && sy >= 0.0f && sx < (float)width && sy < (float)height && opacity > 0.0f)) return 0;
    *pixel = (u32)sy * width + (u32)sx;
    float zc = fminf(fmaxf(z, 0.0f), 1.0f);
    float idx = zc * 255.0f;
    float mix = idx - truncf(idx);
    u32 lo = pal[(u32)floorf(idx)], hi = pal[(u32)ceilf(idx)];
    float lr = (float)(lo & 0xFFu), lg = (float)((lo >> 8) & 0xFFu), lb = (float)((lo >> 16) & 0xFFu);
    float hr = (float)(hi & 0xFFu), hg = (float)((hi >> 8) & 0xFFu), hb = (float)((hi >> 16) & 0xFFu);
    float r = fmaf(mix, hr - lr, lr), g = fmaf(mix, hg - lg, lg), b = fmaf(mix, hb - lb, lb);
    inc[0] = (u32)fmaf(r, opacity, 0.5f);
    inc[1] = (u32)fmaf(g, opacity, 0.5f);
    inc[2] = (u32)fmaf(b, opacity, 0.5f);
    inc[3] = (u32)fmaf(opacity, 255.0f, 0.5f);
    return 1;
}

void orc_bin_points(const float* S, const u32* pal, const float* pts, size_t n, u32* hist, u32 width, u32 height) {
    for (size_t i = 0; i < n; i++) {
        u32 px, inc[4];
        if (make_draw(S, pal, width, height, pts[4 * i], pts[4 * i + 1], pts[4 * i + 2], pts[4 * i + 3], &px, inc)) {
            for (int k = 0; k < 4; k++) hist[4ull * px + k] += inc[k];
        }
    }
}

/* --------------------------------------------------------- block loop */
/* Runs `npasses` passes of one block (tpb walkers) in lockstep.
 * x/y/c: the block's walkers (in/out). If gpu_trace is given (one-step mode),
 * the inputs of pass i > 0 are the GPU's outputs of pass i-1 (permuted as the
 * GPU permutes them), so only one step of error is compared per pass.
 * trace: per pass per walker (x, y, c, opacity); xids: per pass per wave.
 * hist: draw into it (NULL: no drawing). Returns the number of draws. */
/* Xaos flames: every walker picks its own xform from the row of its last
 * one (lane stream: the pick is the pass's first draw), walkers keep their
 * own point (no permutation); xids per pass per walker. */
static u64 run_block_chaos(const flame_prog* fp, const float* P, const u32* pal, u32 block, u32 tpb, u32 seed,
                           u32 stream, u32 pass0, u32 npasses, real* x, real* y, real* c, u32* last,
                           real* trace, u32* xids, const float* gpu_trace, const u32* gpu_xids, u32 trace_stride,
                           u32* hist, u32 width, u32 height) {
    u64 draws = 0;
    float S[6];
    for (int k = 0; k < 6; k++) S[k] = P[k];
    for (u32 i = 0; i < npasses; i++) {
        u32 pass = pass0 + i;
        for (u32 t = 0; t < tpb; t++) {
            size_t g = (size_t)block * tpb + t;
            if (gpu_trace && i > 0) {
                const float* gt = gpu_trace + ((size_t)(i - 1) * trace_stride + g) * 4;
                x[t] = gt[0];
                y[t] = gt[1];
                c[t] = gt[2];
                last[t] = gpu_xids[(size_t)(i - 1) * trace_stride + g];
            }
            lane_rng rs;
            rng_init(&rs, seed, stream, (u32)g, TAG_LANE);
            rng_begin(&rs, pass);
            u32 xid = select_xform_chaos(fp, P, last[t], (float)rand01(&rs));
            if (xids) xids[(size_t)i * trace_stride + g] = xid;
            real px = x[t], py = y[t], pz = c[t];
            real op = apply_xform(fp, (int)xid, P, &px, &py, &pz, &rs);
            if (badvalue((float)px) || badvalue((float)py)) {
                px = rand01(&rs) * 2.0 - 1.0;
                py = rand01(&rs) * 2.0 - 1.0;
                op = 0.0;
            }
            if (hist) {
                real dx = px, dy = py, dz = pz;
                if (fp->has_final) apply_xform(fp, fp->n_xforms, P, &dx, &dy, &dz, &rs);
                u32 pix, inc[4];
                if (make_draw(S, pal, width, height, (float)dx, (float)dy, (float)dz, (float)op, &pix, inc)) {
                    for (int k = 0; k < 4; k++) hist[4ull * pix + k] += inc[k];
                    draws++;
                }
            }
            if (trace) {
                real* tr = trace + ((size_t)i * trace_stride + g) * 4;
                tr[0] = px; tr[1] = py; tr[2] = pz; tr[3] = op;
            }
            x[t] = px; y[t] = py; c[t] = pz;
            last[t] = xid;
        }
    }
    return draws;
}

static u64 run_block(const flame_prog* fp, const float* P, const u32* pal, u32 block, u32 tpb, u32 tpb_bits,
                     u32 seed, u32 stream, u32 pass0, u32 npasses, real* x, real* y, real* c, real* trace,
                     u32* xids, const float* gpu_trace, u32 trace_stride, u32* hist, u32 width, u32 height,
                     real* ox, real* oy, real* oc) {
    u64 draws = 0;
    float S[6];
    for (int k = 0; k < 6; k++) S[k] = P[k];
    u32 key[2] = {seed, stream};
    for (u32 i = 0; i < npasses; i++) {
        u32 pass = pass0 + i;
        u32 pick[4], permw[4];
        u32 nwaves = (tpb + 63) / 64;
        u32 pctr[4] = {pass >> 1, block, 0u, TAG_PERM};
        orc_philox(pctr, key, permw);
        u32 wa = (pass & 1u) ? permw[2] : permw[0], wb = (pass & 1u) ? permw[3] : permw[1];
        if (gpu_trace && i > 0) {
            /* inputs from the GPU's previous outputs, permuted */
            u32 pp = pass - 1;
            u32 pctr2[4] = {pp >> 1, block, 0u, TAG_PERM}, pw[4];
            orc_philox(pctr2, key, pw);
            u32 pwa = (pp & 1u) ? pw[2] : pw[0], pwb = (pp & 1u) ? pw[3] : pw[1];
            for (u32 t = 0; t < tpb; t++) {
                const float* g = gpu_trace + ((size_t)(i - 1) * trace_stride + (size_t)block * tpb + t) * 4;
                u32 slot = permute(t, pwa, pwb, tpb_bits);
                x[slot] = g[0];
                y[slot] = g[1];
                c[slot] = g[2];
            }
        }
        for (u32 wv = 0; wv < nwaves; wv++) {
            u32 ctr[4] = {pass >> 2, block, wv, TAG_PICK};
            orc_philox(ctr, key, pick);
            u32 word = pick[pass & 3u];
            u32 xid = select_xform(fp, P, u01(word));
            if (xids) xids[(size_t)i * (trace_stride / 64) + (size_t)block * (tpb / 64) + wv] = xid;
            u32 t_end = (wv + 1) * 64 < tpb ? (wv + 1) * 64 : tpb;
            for (u32 t = wv * 64; t < t_end; t++) {
                lane_rng rs;
                rng_init(&rs, seed, stream, block * tpb + t, TAG_LANE);
                rng_begin(&rs, pass);
                real px = x[t], py = y[t], pz = c[t];
                real op = apply_xform(fp, (int)xid, P, &px, &py, &pz, &rs);
                if (badvalue((float)px) || badvalue((float)py)) {
                    px = rand01(&rs) * 2.0 - 1.0;
                    py = rand01(&rs) * 2.0 - 1.0;
                    op = 0.0;
                }
                if (hist) {
                    real dx = px, dy = py, dz = pz;
                    if (fp->has_final) apply_xform(fp, fp->n_xforms, P, &dx, &dy, &dz, &rs);
                    u32 pix, inc[4];
                    if (make_draw(S, pal, width, height, (float)dx, (float)dy, (float)dz, (float)op, &pix, inc)) {
                        for (int k = 0; k < 4; k++) hist[4ull * pix + k] += inc[k];
                        draws++;
                    }
                }
                if (trace) {
                    real* tr = trace + ((size_t)i * trace_stride + (size_t)block * tpb + t) * 4;
                    tr[0] = px; tr[1] = py; tr[2] = pz; tr[3] = op;
                }
                ox[t] = px; oy[t] = py; oc[t] = pz;
            }
        }
        for (u32 t = 0; t < tpb; t++) {
            u32 slot = permute(t, wa, wb, tpb_bits);
            x[slot] = ox[t];
            y[slot] = oy[t];
            c[slot] = oc[t];
        }
    }
    return draws;
}

static u32 bits_of(u32 tpb) {
    u32 b = 0;
    while ((1u << b) < tpb) b++;
    return b;
}

/* Trace `npasses` passes for every block (see run_block). x/y/c: all walkers. */
int orc_trace(const int* prog, int n_xforms, int has_final, int chaos_base, const float* params, u32 stride,
              u32 blocks, u32 tpb, u32 seed, u32 stream, u32 pass0, u32 npasses, real* x, real* y, real* c,
              u32* last, real* trace, u32* xids, const float* gpu_trace, const u32* gpu_xids) {
    flame_prog fp;
    prog_index(&fp, prog, n_xforms, has_final, chaos_base);
    u32 bits = bits_of(tpb);
    int ok = 1;
    if (chaos_base >= 0) {
        if (!last) return -1;
#pragma omp parallel for schedule(dynamic)
        for (long b = 0; b < (long)blocks; b++)
            run_block_chaos(&fp, params + (size_t)b * stride, NULL, (u32)b, tpb, seed, stream, pass0, npasses,
                            x + (size_t)b * tpb, y + (size_t)b * tpb, c + (size_t)b * tpb, last + (size_t)b * tpb,
                            trace, xids, gpu_trace, gpu_xids, blocks * tpb, NULL, 0, 0);
        return 0;
    }
#pragma omp parallel for schedule(dynamic)
    for (long b = 0; b < (long)blocks; b++) {
        real* ox = (real*)malloc(sizeof(real) * tpb * 3);
        if (!ox) { ok = 0; continue; }
        run_block(&fp, params + (size_t)b * stride, NULL, (u32)b, tpb, bits, seed, stream, pass0, npasses,
                  x + (size_t)b * tpb, y + (size_t)b * tpb, c + (size_t)b * tpb, trace, xids, gpu_trace,
                  blocks * tpb, NULL, 0, 0, ox, ox + tpb, ox + 2 * tpb);
        free(ox);
    }
    return ok ? 0 : -1;
}

/* Walker seeding: Hammersley point of gid over `total` mapped to pixels and
 * through plane_space (params[6..11]); colour from the INIT stream. */
void orc_seed(const float* params, u32 stride, u32 blocks, u32 tpb, u32 width, u32 height, u32 seed, u32 stream,
              real* x, real* y, real* c) {
    u32 total = blocks * tpb;
    u32 max_val = total;
    if (total % 2u) {
        u32 v = total - 1u;
        v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
        max_val = v + 1u;
    }
    u32 radix = 0, v = max_val;
    while (v >>= 1) radix++;
    for (u32 g = 0; g < total; g++) {
        const float* P = params + (size_t)(g / tpb) * stride;
        u32 rev = 0;
        for (int k = 0; k < 32; k++) if (g & (1u << k)) rev |= 1u << (31 - k);
        real hx = (real)g / max_val * 2.0 - 1.0;
        real hy = (radix ? (real)(rev >> (32u - radix)) : 0.0) / max_val * 2.0 - 1.0;
        real px = (hx + 1.0) / 2.0 * width, py = (hy + 1.0) / 2.0 * height;
        x[g] = (real)P[6] * px + (real)P[8] * py + P[10];
        y[g] = (real)P[7] * px + (real)P[9] * py + P[11];
        u32 ctr[4] = {0u, g, 0u, TAG_INIT}, key[2] = {seed, stream}, out[4];
        orc_philox(ctr, key, out);
        c[g] = (real)u01(out[0]);
    }
}

/* Full render on host cores: seed, fuse passes, then `passes` drawn passes,
 * per-thread private histograms summed at the end. Returns total draws. */
u64 orc_render(const int* prog, int n_xforms, int has_final, int chaos_base, const float* params, u32 stride,
               const u32* palettes, u32 blocks, u32 tpb, u32 width, u32 height, u32 seed, u32 stream, u32 fuse,
               u32 passes, u32* hist, int threads) {
    flame_prog fp;
    prog_index(&fp, prog, n_xforms, has_final, chaos_base);
    u32 bits = bits_of(tpb);
    size_t total = (size_t)blocks * tpb, npx = (size_t)width * height;
    real* x = (real*)malloc(sizeof(real) * total);
    real* y = (real*)malloc(sizeof(real) * total);
    real* c = (real*)malloc(sizeof(real) * total);
    if (!x || !y || !c) { free(x); free(y); free(c); return 0; }
    orc_seed(params, stride, blocks, tpb, width, height, seed, stream, x, y, c);
    u64 draws = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel reduction(+ : draws)
    {
        u32* local = (u32*)calloc(npx * 4, sizeof(u32));
        real* ox = (real*)malloc(sizeof(real) * tpb * 3);
#pragma omp for schedule(dynamic)
        for (long b = 0; b < (long)blocks; b++) {
            const float* P = params + (size_t)b * stride;
            real *bx = x + (size_t)b * tpb, *by = y + (size_t)b * tpb, *bc = c + (size_t)b * tpb;
            if (chaos_base >= 0) {
                u32* last = (u32*)malloc(sizeof(u32) * tpb);
                for (u32 t = 0; t < tpb; t++) last[t] = 0xFFFFFFFFu;
                run_block_chaos(&fp, P, NULL, (u32)b, tpb, seed, stream, 0, fuse, bx, by, bc, last, NULL, NULL, NULL,
                                NULL, 0, NULL, 0, 0);
                draws += run_block_chaos(&fp, P, palettes + (size_t)b * 256, (u32)b, tpb, seed, stream, fuse, passes,
                                         bx, by, bc, last, NULL, NULL, NULL, NULL, 0, local, width, height);
                free(last);
                continue;
            }
            run_block(&fp, P, NULL, (u32)b, tpb, bits, seed, stream, 0, fuse, bx, by, bc, NULL, NULL, NULL, 0, NULL,
                      0, 0, ox, ox + tpb, ox + 2 * tpb);
            draws += run_block(&fp, P, palettes + (size_t)b * 256, (u32)b, tpb, bits, seed, stream, fuse, passes, bx,
                               by, bc, NULL, NULL, NULL, 0, local, width, height, ox, ox + tpb, ox + 2 * tpb);
        }
#pragma omp critical
        {
            for (size_t i = 0; i < npx * 4; i++) hist[i] += local[i];
        }
        free(local);
        free(ox);
    }
    free(x); free(y); free(c);
    return draws;
}

int orc_variation_count(void) { return V_COUNT; }
