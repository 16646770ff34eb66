// Learned denoiser behind the C ABI: a 4-level U-Net (3x3 convolutions, ReLU,
// 2x2 average pooling, nearest upsampling, skip concatenation) predicting an
// RGB residual over the tonemapped low-iteration frame. Replaces the
// reference's third-party denoisers behind rfkt::denoiser
// (src/librefrakt/include/librefrakt/interface/denoiser.h:21-39); training runs
// on the GPU through the same kernels (forward, data and weight gradients,
// Adam), so a frame renders with a fraction of the iterations.
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/refrakt.h"
#include "unet_kernels.hpp"

namespace rfk {
int set_error(int code, const std::string& msg);
}

namespace {

int fail(int code, const std::string& m) { return rfk::set_error(code, m); }
#define NN_HIP(expr)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(RFK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kFeat = 8;   // input channels (6 features + padding)
constexpr int kOut = 8;    // output channels (3 used: RGB residual)
constexpr uint32_t kMagic = 0x31544E55;  // "UNT1"

struct conv_def {
    int c1, c2, cout, act;
    size_t w_off, b_off;  // offsets (elements) into the weight / bias arrays
};

struct buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct rfk_denoiser {
    int ch[4] = {32, 64, 96, 128};
    std::vector<conv_def> convs;
    size_t nw = 0, nb = 0;
    // parameters: FP32 master weights and biases, BF16 copy, Adam moments, grads
    float *w = nullptr, *b = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    float *gw = nullptr, *gb = nullptr;
    void* wbf = nullptr;   // BF16 weights
    void* wflip = nullptr; // BF16 flipped weights (largest conv)
    int step = 0;
    // activation buffers per (N, H, W)
    std::map<std::tuple<int, int, int>, std::map<std::string, buf>> acts;
    float* loss_acc = nullptr;
    float* wspace = nullptr;
    size_t wspace_bytes = 0;

    ~rfk_denoiser() {
        for (void* p : {(void*)w, (void*)b, (void*)mw, (void*)vw, (void*)mb, (void*)vb, (void*)gw, (void*)gb, wbf, wflip,
                        (void*)loss_acc, (void*)wspace})
            if (p) (void)hipFree(p);
        for (auto& [k, m] : acts)
            for (auto& [n, bb] : m)
                if (bb.p) (void)hipFree(bb.p);
    }
};

namespace {

void define_net(rfk_denoiser* d) {
    const int c0 = d->ch[0], c1 = d->ch[1], c2 = d->ch[2], c3 = d->ch[3];
    struct L { int a, b, o, act; };
    const L layers[] = {
        {kFeat, 0, c0, 1}, {c0, 0, c0, 1},       // enc0
        {c0, 0, c1, 1},    {c1, 0, c1, 1},       // enc1
        {c1, 0, c2, 1},    {c2, 0, c2, 1},       // enc2
        {c2, 0, c3, 1},    {c3, 0, c3, 1},       // bottleneck
        {c3, c2, c2, 1},   {c2, 0, c2, 1},       // dec2: cat(up(b), e2)
        {c2, c1, c1, 1},   {c1, 0, c1, 1},       // dec1: cat(up(d2), e1)
        {c1, c0, c0, 1},   {c0, 0, c0, 1},       // dec0: cat(up(d1), e0)
        {c0, 0, kOut, 0},                        // output residual
    };
    d->convs.clear();
    size_t wo = 0, bo = 0;
    for (auto& l : layers) {
        d->convs.push_back({l.a, l.b, l.o, l.act, wo, bo});
        wo += size_t(l.o) * 9 * (l.a + l.b);
        bo += size_t(l.o);
    }
    d->nw = wo;
    d->nb = bo;
}

uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int alloc_params(rfk_denoiser* d) {
    auto A = [&](float** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(float) + 64); };
    NN_HIP(A(&d->w, d->nw));
    NN_HIP(A(&d->b, d->nb));
    NN_HIP(A(&d->mw, d->nw));
    NN_HIP(A(&d->vw, d->nw));
    NN_HIP(A(&d->mb, d->nb));
    NN_HIP(A(&d->vb, d->nb));
    NN_HIP(A(&d->gw, d->nw));
    NN_HIP(A(&d->gb, d->nb));
    NN_HIP(hipMalloc(&d->wbf, d->nw * 2 + 64));
    size_t maxw = 0;
    for (auto& c : d->convs) maxw = std::max(maxw, size_t(c.cout) * 9 * (c.c1 + c.c2));
    NN_HIP(hipMalloc(&d->wflip, maxw * 2 + 64));
    NN_HIP(hipMalloc(reinterpret_cast<void**>(&d->loss_acc), 64));
    NN_HIP(hipMemset(d->mw, 0, d->nw * 4));
    NN_HIP(hipMemset(d->vw, 0, d->nw * 4));
    NN_HIP(hipMemset(d->mb, 0, d->nb * 4));
    NN_HIP(hipMemset(d->vb, 0, d->nb * 4));
    return RFK_OK;
}

// New parameters restart the optimiser: Adam's moments are zeroed and its step
// count reset, so the bias correction of the first steps applies again (the
// moments are not part of the weight file).
int upload_weights(rfk_denoiser* d, const std::vector<float>& w, const std::vector<float>& b) {
    NN_HIP(hipMemcpy(d->w, w.data(), w.size() * 4, hipMemcpyHostToDevice));
    NN_HIP(hipMemcpy(d->b, b.data(), b.size() * 4, hipMemcpyHostToDevice));
    NN_HIP(hipMemset(d->mw, 0, d->nw * 4));
    NN_HIP(hipMemset(d->vw, 0, d->nw * 4));
    NN_HIP(hipMemset(d->mb, 0, d->nb * 4));
    NN_HIP(hipMemset(d->vb, 0, d->nb * 4));
    d->step = 0;
    NN_HIP(rfk::nn::f32_to_bf16(d->w, d->wbf, d->nw, nullptr));
    NN_HIP(hipDeviceSynchronize());
    return RFK_OK;
}

// He-normal initialisation from a SplitMix64 stream (Box-Muller)
void init_weights(rfk_denoiser* d, uint64_t seed, std::vector<float>& w, std::vector<float>& b) {
    w.assign(d->nw, 0.f);
    b.assign(d->nb, 0.f);
    uint64_t s = seed;
    for (size_t li = 0; li < d->convs.size(); li++) {
        auto& c = d->convs[li];
        const size_t n = size_t(c.cout) * 9 * (c.c1 + c.c2);
        double sd = std::sqrt(2.0 / (9.0 * (c.c1 + c.c2)));
        if (li + 1 == d->convs.size()) sd *= 0.1;  // start close to the identity residual
        for (size_t i = 0; i < n; i++) {
            double u1 = (double(splitmix(s) >> 11) + 1.0) / 9007199254740993.0;
            double u2 = double(splitmix(s) >> 11) / 9007199254740992.0;
            w[c.w_off + i] = float(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
    }
}

// Allocations that the next act_buf calls refuse without calling hipMalloc
// (rfk_nn_inject_alloc_failures: error-path tests).
std::atomic<int> g_inject_alloc_failures{0};

// The activation buffer `name` for shape (N, H, W), grown to `bytes`. On
// failure rc is set (and never cleared here: callers check once after a run
// of calls), the returned buffer is null and its recorded size stays 0 so the
// next call tries again.
buf& act_buf(rfk_denoiser* d, int N, int H, int W, const std::string& name, size_t bytes, int& rc) {
    auto& m = d->acts[{N, H, W}];
    buf& b = m[name];
    if (b.bytes < bytes || !b.p) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        int left = g_inject_alloc_failures.load();
        while (left > 0 && !g_inject_alloc_failures.compare_exchange_weak(left, left - 1)) {}
        if (left > 0 || hipMalloc(&b.p, bytes + 256) != hipSuccess) {
            b.p = nullptr;
            rc = fail(RFK_ERR_HIP, "activation allocation failed");
        } else {
            b.bytes = bytes;
        }
    }
    return b;
}

struct level_shape {
    int H, W;
};

// Forward pass; activations are kept under their names for the backward pass.
// fused (inference): upsamples are read through by the next convolution
// instead of being stored, and the average pools are written by the
// convolutions that produce their inputs; training keeps every activation
// for the backward pass.
// With out_rgb (inference only), the output layer writes clamp(base + r)
// directly (rout untouched) where the strip kernel takes its shape; the
// return value 1 tells the caller that out_rgb is done.
int forward(rfk_denoiser* d, int N, int H, int W, const void* feat, float* rout, hipStream_t s, bool fused = false,
            const float* base = nullptr, float* out_rgb = nullptr, bool* rgb_done = nullptr) {
    int rc = 0;
    const level_shape L[4] = {{H, W}, {H / 2, W / 2}, {H / 4, W / 4}, {H / 8, W / 8}};
    auto A = [&](const std::string& name, int lvl, int C) -> void* {
        buf& b = act_buf(d, N, H, W, name, size_t(N) * L[lvl].H * L[lvl].W * C * 2, rc);
        return b.p;
    };
    auto conv = [&](int li, const void* x1, const void* x2, int lvl, void* y, float* yf, int mode = 0,
                    void* ypool = nullptr) -> int {
        const conv_def& c = d->convs[size_t(li)];
        NN_HIP(rfk::nn::conv3x3_forward(x1, c.c1, x2, c.c2, N, L[lvl].H, L[lvl].W,
                                        static_cast<char*>(d->wbf) + c.w_off * 2, d->b + c.b_off, c.cout, c.act, 0.f,
                                        y, yf, s, mode, ypool));
        return RFK_OK;
    };
    const int* ch = d->ch;
    if (fused) {
        void *a0 = A("a0", 0, ch[0]), *e0 = A("e0", 0, ch[0]), *p0 = A("p0", 1, ch[0]);
        void *a1 = A("a1", 1, ch[1]), *e1 = A("e1", 1, ch[1]), *p1 = A("p1", 2, ch[1]);
        void *a2 = A("a2", 2, ch[2]), *e2 = A("e2", 2, ch[2]), *p2 = A("p2", 3, ch[2]);
        void *a3 = A("a3", 3, ch[3]), *bb = A("b", 3, ch[3]);
        void *c2 = A("c2", 2, ch[2]), *d2 = A("d2", 2, ch[2]);
        void *c1 = A("c1", 1, ch[1]), *d1 = A("d1", 1, ch[1]);
        void *c0 = A("c0", 0, ch[0]), *d0 = A("d0", 0, ch[0]);
        if (rc) return rc;
        // the encoder's average pools come out of the convolutions' epilogues
        // features -> c0 -> c1 (+ pool) in one pass where the shapes allow
        {
            const conv_def &ca = d->convs[0], &cb = d->convs[1];
            const hipError_t e = ca.c2 == 0 && cb.c2 == 0 && cb.c1 == ca.cout
                                     ? rfk::nn::conv3x3_pair_forward(
                                           feat, ca.c1, N, H, W, static_cast<char*>(d->wbf) + ca.w_off * 2,
                                           d->b + ca.b_off, ca.act, static_cast<char*>(d->wbf) + cb.w_off * 2,
                                           d->b + cb.b_off, cb.cout, cb.act, e0, p0, nullptr, nullptr, s)
                                     : hipErrorInvalidValue;
            if (e == hipErrorInvalidValue) {
                (void)hipGetLastError();
                if ((rc = conv(0, feat, nullptr, 0, a0, nullptr))) return rc;
                if ((rc = conv(1, a0, nullptr, 0, e0, nullptr, 0, p0))) return rc;
            } else if (e != hipSuccess) {
                return fail(RFK_ERR_HIP, std::string("first level: ") + hipGetErrorString(e));
            }
        }
        if ((rc = conv(2, p0, nullptr, 1, a1, nullptr))) return rc;
        if ((rc = conv(3, a1, nullptr, 1, e1, nullptr, 0, p1))) return rc;
        if ((rc = conv(4, p1, nullptr, 2, a2, nullptr))) return rc;
        if ((rc = conv(5, a2, nullptr, 2, e2, nullptr, 0, p2))) return rc;
        if ((rc = conv(6, p2, nullptr, 3, a3, nullptr))) return rc;
        if ((rc = conv(7, a3, nullptr, 3, bb, nullptr))) return rc;
        if ((rc = conv(8, bb, e2, 2, c2, nullptr, 1))) return rc;
        if ((rc = conv(9, c2, nullptr, 2, d2, nullptr))) return rc;
        if ((rc = conv(10, d2, e1, 1, c1, nullptr, 1))) return rc;
        if ((rc = conv(11, c1, nullptr, 1, d1, nullptr))) return rc;
        if ((rc = conv(12, d1, e0, 0, c0, nullptr, 1))) return rc;
        if (out_rgb) {
            // c13 -> output layer -> rgb in one pass where the shapes allow
            const conv_def &ca = d->convs[13], &cb = d->convs[14];
            const hipError_t e = ca.c2 == 0 && cb.c2 == 0 && cb.c1 == ca.cout
                                     ? rfk::nn::conv3x3_pair_forward(
                                           c0, ca.c1, N, H, W, static_cast<char*>(d->wbf) + ca.w_off * 2,
                                           d->b + ca.b_off, ca.act, static_cast<char*>(d->wbf) + cb.w_off * 2,
                                           d->b + cb.b_off, cb.cout, cb.act, nullptr, nullptr, base, out_rgb, s)
                                     : hipErrorInvalidValue;
            if (e == hipSuccess) {
                if (rgb_done) *rgb_done = true;
                return RFK_OK;
            }
            if (e != hipErrorInvalidValue) return fail(RFK_ERR_HIP, std::string("last level: ") + hipGetErrorString(e));
            (void)hipGetLastError();
        }
        if ((rc = conv(13, c0, nullptr, 0, d0, nullptr))) return rc;
        if (out_rgb) {
            const conv_def& c = d->convs[14];
            const hipError_t e = rfk::nn::conv3x3_forward_rgb(d0, c.c1, N, H, W, static_cast<char*>(d->wbf) + c.w_off * 2,
                                                              d->b + c.b_off, c.cout, c.act, base, out_rgb, s);
            if (e == hipSuccess) {
                if (rgb_done) *rgb_done = true;
                return RFK_OK;
            }
            // only an unsupported shape (refused before any launch) falls back
            // to the plain output layer; every other error is the caller's
            if (e != hipErrorInvalidValue) return fail(RFK_ERR_HIP, std::string("output layer: ") + hipGetErrorString(e));
            (void)hipGetLastError();
        }
        return conv(14, d0, nullptr, 0, nullptr, rout);
    }
    void *a0 = A("a0", 0, ch[0]), *e0 = A("e0", 0, ch[0]), *p0 = A("p0", 1, ch[0]);
    void *a1 = A("a1", 1, ch[1]), *e1 = A("e1", 1, ch[1]), *p1 = A("p1", 2, ch[1]);
    void *a2 = A("a2", 2, ch[2]), *e2 = A("e2", 2, ch[2]), *p2 = A("p2", 3, ch[2]);
    void *a3 = A("a3", 3, ch[3]), *bb = A("b", 3, ch[3]);
    void *u2 = A("u2", 2, ch[3]), *c2 = A("c2", 2, ch[2]), *d2 = A("d2", 2, ch[2]);
    void *u1 = A("u1", 1, ch[2]), *c1 = A("c1", 1, ch[1]), *d1 = A("d1", 1, ch[1]);
    void *u0 = A("u0", 0, ch[1]), *c0 = A("c0", 0, ch[0]), *d0 = A("d0", 0, ch[0]);
    if (rc) return rc;
    if ((rc = conv(0, feat, nullptr, 0, a0, nullptr))) return rc;
    if ((rc = conv(1, a0, nullptr, 0, e0, nullptr))) return rc;
    NN_HIP(rfk::nn::avgpool2_forward(e0, N, H, W, ch[0], p0, s));
    if ((rc = conv(2, p0, nullptr, 1, a1, nullptr))) return rc;
    if ((rc = conv(3, a1, nullptr, 1, e1, nullptr))) return rc;
    NN_HIP(rfk::nn::avgpool2_forward(e1, N, L[1].H, L[1].W, ch[1], p1, s));
    if ((rc = conv(4, p1, nullptr, 2, a2, nullptr))) return rc;
    if ((rc = conv(5, a2, nullptr, 2, e2, nullptr))) return rc;
    NN_HIP(rfk::nn::avgpool2_forward(e2, N, L[2].H, L[2].W, ch[2], p2, s));
    if ((rc = conv(6, p2, nullptr, 3, a3, nullptr))) return rc;
    if ((rc = conv(7, a3, nullptr, 3, bb, nullptr))) return rc;
    NN_HIP(rfk::nn::upsample2_forward(bb, N, L[3].H, L[3].W, ch[3], u2, s));
    if ((rc = conv(8, u2, e2, 2, c2, nullptr))) return rc;
    if ((rc = conv(9, c2, nullptr, 2, d2, nullptr))) return rc;
    NN_HIP(rfk::nn::upsample2_forward(d2, N, L[2].H, L[2].W, ch[2], u1, s));
    if ((rc = conv(10, u1, e1, 1, c1, nullptr))) return rc;
    if ((rc = conv(11, c1, nullptr, 1, d1, nullptr))) return rc;
    NN_HIP(rfk::nn::upsample2_forward(d1, N, L[1].H, L[1].W, ch[1], u0, s));
    if ((rc = conv(12, u0, e0, 0, c0, nullptr))) return rc;
    if ((rc = conv(13, c0, nullptr, 0, d0, nullptr))) return rc;
    if ((rc = conv(14, d0, nullptr, 0, nullptr, rout))) return rc;
    return RFK_OK;
}

int ensure_workspace(rfk_denoiser* d, size_t bytes) {
    if (d->wspace_bytes >= bytes) return RFK_OK;
    if (d->wspace) (void)hipFree(d->wspace);
    d->wspace = nullptr;
    NN_HIP(hipMalloc(reinterpret_cast<void**>(&d->wspace), bytes + 256));
    d->wspace_bytes = bytes;
    return RFK_OK;
}

// Backward pass from dR (BF16, kOut channels at full resolution); fills gw/gb.
int backward(rfk_denoiser* d, int N, int H, int W, const void* feat, const void* dR, hipStream_t s) {
    int rc = 0;
    const level_shape L[4] = {{H, W}, {H / 2, W / 2}, {H / 4, W / 4}, {H / 8, W / 8}};
    auto& M = d->acts[{N, H, W}];
    auto get = [&](const char* n) { return M[n].p; };
    auto G = [&](const std::string& name, int lvl, int C) -> void* {
        return act_buf(d, N, H, W, "g_" + name, size_t(N) * L[lvl].H * L[lvl].W * C * 2, rc).p;
    };
    const int* ch = d->ch;
    size_t maxc = 0;
    for (auto& c : d->convs) maxc = std::max(maxc, size_t(c.c1 + c.c2));
    void* gt = act_buf(d, N, H, W, "g_tmp", size_t(N) * H * W * maxc * 2, rc).p;    // dgrad output
    void* gt2 = act_buf(d, N, H, W, "g_tmp2", size_t(N) * H * W * maxc * 2, rc).p;  // act-grad output
    if (rc) return rc;
    size_t ws = 0;
    for (int li = 0; li < int(d->convs.size()); li++) {
        auto& c = d->convs[size_t(li)];
        int lvl = li < 2 ? 0 : li < 4 ? 1 : li < 6 ? 2 : li < 8 ? 3 : li < 10 ? 2 : li < 12 ? 1 : 0;
        ws = std::max(ws, rfk::nn::conv3x3_wgrad_workspace(c.c1, c.c2, c.cout, N, L[lvl].H, L[lvl].W, nullptr, nullptr));
    }
    if ((rc = ensure_workspace(d, ws))) return rc;

    // one conv layer backward: gy is the gradient w.r.t. the layer output
    // (after activation); y its output; x1/x2 its inputs. Writes dgrad into gt.
    auto layer = [&](int li, int lvl, const void* y, const void* gy, const void* x1, const void* x2,
                     bool need_dx) -> int {
        const conv_def& c = d->convs[size_t(li)];
        const int h = L[lvl].H, w = L[lvl].W;
        const size_t npix = size_t(N) * h * w;
        const void* gpre = gy;
        if (c.act) {
            NN_HIP(rfk::nn::act_backward(y, gy, npix * c.cout, c.act, 0.f, gt2, s));
            gpre = gt2;
        }
        NN_HIP(rfk::nn::conv3x3_wgrad(x1, c.c1, x2, c.c2, gpre, c.cout, N, h, w, d->gw + c.w_off, 0.f, d->wspace, s));
        NN_HIP(rfk::nn::bias_grad(gpre, npix, c.cout, d->gb + c.b_off, 0.f, s));
        if (need_dx) {
            NN_HIP(rfk::nn::flip_weights(static_cast<char*>(d->wbf) + c.w_off * 2, c.cout, c.c1 + c.c2, d->wflip, s));
            NN_HIP(rfk::nn::conv3x3_forward(gpre, c.cout, nullptr, 0, N, h, w, d->wflip, nullptr, c.c1 + c.c2, 0, 0.f,
                                            gt, nullptr, s));
        }
        return RFK_OK;
    };
    auto npx = [&](int lvl) { return size_t(N) * L[lvl].H * L[lvl].W; };
    void *g_e0 = G("e0", 0, ch[0]), *g_e1 = G("e1", 1, ch[1]), *g_e2 = G("e2", 2, ch[2]);
    void *g_d0 = G("d0", 0, ch[0]), *g_c0 = G("c0", 0, ch[0]), *g_u0 = G("u0", 0, ch[1]);
    void *g_d1 = G("d1", 1, ch[1]), *g_c1 = G("c1", 1, ch[1]), *g_u1 = G("u1", 1, ch[2]);
    void *g_d2 = G("d2", 2, ch[2]), *g_c2 = G("c2", 2, ch[2]), *g_u2 = G("u2", 2, ch[3]);
    void *g_b = G("b", 3, ch[3]), *g_a3 = G("a3", 3, ch[3]), *g_p2 = G("p2", 3, ch[2]);
    void *g_a2 = G("a2", 2, ch[2]), *g_p1 = G("p1", 2, ch[1]), *g_a1 = G("a1", 1, ch[1]), *g_p0 = G("p0", 1, ch[0]);
    void* g_a0 = G("a0", 0, ch[0]);
    void* g_pool = G("pool", 0, ch[0] > ch[1] ? ch[0] : ch[1]);
    if (rc) return rc;
    NN_HIP(hipMemsetAsync(g_e0, 0, npx(0) * ch[0] * 2, s));
    NN_HIP(hipMemsetAsync(g_e1, 0, npx(1) * ch[1] * 2, s));
    NN_HIP(hipMemsetAsync(g_e2, 0, npx(2) * ch[2] * 2, s));
    auto copy = [&](void* dst, const void* src, size_t bytes) { return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s); };

    // output conv (no activation)
    if ((rc = layer(14, 0, nullptr, dR, get("d0"), nullptr, true))) return rc;
    NN_HIP(copy(g_d0, gt, npx(0) * ch[0] * 2));
    if ((rc = layer(13, 0, get("d0"), g_d0, get("c0"), nullptr, true))) return rc;
    NN_HIP(copy(g_c0, gt, npx(0) * ch[0] * 2));
    if ((rc = layer(12, 0, get("c0"), g_c0, get("u0"), get("e0"), true))) return rc;
    NN_HIP(rfk::nn::split_add(gt, npx(0), ch[1], ch[0], g_u0, g_e0, s));
    NN_HIP(rfk::nn::upsample2_backward(g_u0, N, L[1].H, L[1].W, ch[1], g_d1, s));
    if ((rc = layer(11, 1, get("d1"), g_d1, get("c1"), nullptr, true))) return rc;
    NN_HIP(copy(g_c1, gt, npx(1) * ch[1] * 2));
    if ((rc = layer(10, 1, get("c1"), g_c1, get("u1"), get("e1"), true))) return rc;
    NN_HIP(rfk::nn::split_add(gt, npx(1), ch[2], ch[1], g_u1, g_e1, s));
    NN_HIP(rfk::nn::upsample2_backward(g_u1, N, L[2].H, L[2].W, ch[2], g_d2, s));
    if ((rc = layer(9, 2, get("d2"), g_d2, get("c2"), nullptr, true))) return rc;
    NN_HIP(copy(g_c2, gt, npx(2) * ch[2] * 2));
    if ((rc = layer(8, 2, get("c2"), g_c2, get("u2"), get("e2"), true))) return rc;
    NN_HIP(rfk::nn::split_add(gt, npx(2), ch[3], ch[2], g_u2, g_e2, s));
    NN_HIP(rfk::nn::upsample2_backward(g_u2, N, L[3].H, L[3].W, ch[3], g_b, s));
    if ((rc = layer(7, 3, get("b"), g_b, get("a3"), nullptr, true))) return rc;
    NN_HIP(copy(g_a3, gt, npx(3) * ch[3] * 2));
    if ((rc = layer(6, 3, get("a3"), g_a3, get("p2"), nullptr, true))) return rc;
    NN_HIP(copy(g_p2, gt, npx(3) * ch[2] * 2));
    NN_HIP(rfk::nn::avgpool2_backward(g_p2, N, L[2].H, L[2].W, ch[2], g_pool, s));
    NN_HIP(rfk::nn::add_bf16(g_pool, g_e2, npx(2) * ch[2], s));
    if ((rc = layer(5, 2, get("e2"), g_e2, get("a2"), nullptr, true))) return rc;
    NN_HIP(copy(g_a2, gt, npx(2) * ch[2] * 2));
    if ((rc = layer(4, 2, get("a2"), g_a2, get("p1"), nullptr, true))) return rc;
    NN_HIP(copy(g_p1, gt, npx(2) * ch[1] * 2));
    NN_HIP(rfk::nn::avgpool2_backward(g_p1, N, L[1].H, L[1].W, ch[1], g_pool, s));
    NN_HIP(rfk::nn::add_bf16(g_pool, g_e1, npx(1) * ch[1], s));
    if ((rc = layer(3, 1, get("e1"), g_e1, get("a1"), nullptr, true))) return rc;
    NN_HIP(copy(g_a1, gt, npx(1) * ch[1] * 2));
    if ((rc = layer(2, 1, get("a1"), g_a1, get("p0"), nullptr, true))) return rc;
    NN_HIP(copy(g_p0, gt, npx(1) * ch[0] * 2));
    NN_HIP(rfk::nn::avgpool2_backward(g_p0, N, L[0].H, L[0].W, ch[0], g_pool, s));
    NN_HIP(rfk::nn::add_bf16(g_pool, g_e0, npx(0) * ch[0], s));
    if ((rc = layer(1, 0, get("e0"), g_e0, get("a0"), nullptr, true))) return rc;
    NN_HIP(copy(g_a0, gt, npx(0) * ch[0] * 2));
    if ((rc = layer(0, 0, get("a0"), g_a0, feat, nullptr, false))) return rc;
    return RFK_OK;
}

bool shape_ok(int N, int H, int W) { return N > 0 && H > 0 && W > 0 && H % 8 == 0 && W % 8 == 0; }

}  // namespace

extern "C" {

int rfk_denoiser_create(uint64_t seed, rfk_denoiser** out) {
    if (!out) return fail(RFK_ERR_INVALID, "null argument");
    auto d = std::make_unique<rfk_denoiser>();
    define_net(d.get());
    if (int rc = alloc_params(d.get())) return rc;
    std::vector<float> w, b;
    init_weights(d.get(), seed, w, b);
    if (int rc = upload_weights(d.get(), w, b)) return rc;
    *out = d.release();
    return RFK_OK;
}

int rfk_denoiser_free(rfk_denoiser* d) {
    delete d;
    return RFK_OK;
}

int rfk_denoiser_param_count(const rfk_denoiser* d, size_t* weights, size_t* biases) {
    if (!d) return fail(RFK_ERR_INVALID, "null denoiser");
    if (weights) *weights = d->nw;
    if (biases) *biases = d->nb;
    return RFK_OK;
}

int rfk_denoiser_save(const rfk_denoiser* d, const char* path) {
    if (!d || !path) return fail(RFK_ERR_INVALID, "null argument");
    std::vector<float> w(d->nw), b(d->nb);
    NN_HIP(hipMemcpy(w.data(), d->w, d->nw * 4, hipMemcpyDeviceToHost));
    NN_HIP(hipMemcpy(b.data(), d->b, d->nb * 4, hipMemcpyDeviceToHost));
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(RFK_ERR_INVALID, std::string("cannot write ") + path);
    uint32_t hdr[8] = {kMagic, uint32_t(d->ch[0]), uint32_t(d->ch[1]), uint32_t(d->ch[2]), uint32_t(d->ch[3]),
                       uint32_t(d->nw), uint32_t(d->nb), uint32_t(d->step)};
    f.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
    f.write(reinterpret_cast<const char*>(w.data()), std::streamsize(w.size() * 4));
    f.write(reinterpret_cast<const char*>(b.data()), std::streamsize(b.size() * 4));
    return f ? RFK_OK : fail(RFK_ERR_INVALID, "write failed");
}

int rfk_denoiser_load(rfk_denoiser* d, const char* path) {
    if (!d || !path) return fail(RFK_ERR_INVALID, "null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(RFK_ERR_INVALID, std::string("cannot read ") + path);
    uint32_t hdr[8];
    f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    if (!f || hdr[0] != kMagic || hdr[5] != d->nw || hdr[6] != d->nb || int(hdr[1]) != d->ch[0] ||
        int(hdr[4]) != d->ch[3])
        return fail(RFK_ERR_PARSE, "weight file does not match this network");
    std::vector<float> w(d->nw), b(d->nb);
    f.read(reinterpret_cast<char*>(w.data()), std::streamsize(w.size() * 4));
    f.read(reinterpret_cast<char*>(b.data()), std::streamsize(b.size() * 4));
    if (!f) return fail(RFK_ERR_PARSE, "truncated weight file");
    return upload_weights(d, w, b);
}

// weights/biases as FP32 host arrays (layout [conv][Cout][9][Cin] concatenated)
int rfk_denoiser_get_params(const rfk_denoiser* d, float* w, float* b) {
    if (!d) return fail(RFK_ERR_INVALID, "null denoiser");
    if (w) NN_HIP(hipMemcpy(w, d->w, d->nw * 4, hipMemcpyDeviceToHost));
    if (b) NN_HIP(hipMemcpy(b, d->b, d->nb * 4, hipMemcpyDeviceToHost));
    return RFK_OK;
}

int rfk_denoiser_set_params(rfk_denoiser* d, const float* w, const float* b) {
    if (!d || !w || !b) return fail(RFK_ERR_INVALID, "null argument");
    std::vector<float> wv(w, w + d->nw), bv(b, b + d->nb);
    return upload_weights(d, wv, bv);
}

int rfk_denoiser_get_grads(const rfk_denoiser* d, float* gw, float* gb) {
    if (!d) return fail(RFK_ERR_INVALID, "null denoiser");
    if (gw) NN_HIP(hipMemcpy(gw, d->gw, d->nw * 4, hipMemcpyDeviceToHost));
    if (gb) NN_HIP(hipMemcpy(gb, d->gb, d->nb * 4, hipMemcpyDeviceToHost));
    return RFK_OK;
}

// Build the network input from a fixed-point histogram and its tonemapped
// float4 image: features (BF16, N*H*W*8) and the FP32 RGB base (N*H*W*3).
int rfk_denoiser_features(const uint32_t* hist, int hist_bits, const float* tonemapped_rgba, size_t npix, void* feat,
                          float* base, void* stream) {
    if (!hist || !tonemapped_rgba || !feat || !base) return fail(RFK_ERR_INVALID, "null argument");
    if (hist_bits != 32 && hist_bits != 64) return fail(RFK_ERR_INVALID, "hist_bits must be 32 or 64");
    NN_HIP(rfk::nn::features(hist, hist_bits == 64, tonemapped_rgba, npix, kFeat, feat, base,
                             static_cast<hipStream_t>(stream)));
    return RFK_OK;
}

// Inference: out_rgb (FP32, N*H*W*3, clamped to [0,1]) = base + U-Net(feat).
int rfk_denoiser_forward(rfk_denoiser* d, const void* feat, const float* base, int N, int H, int W, float* out_rgb,
                         void* stream) {
    if (!d || !feat || !base || !out_rgb) return fail(RFK_ERR_INVALID, "null argument");
    if (!shape_ok(N, H, W)) return fail(RFK_ERR_INVALID, "height and width must be multiples of 8");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = 0;
    const size_t npix = size_t(N) * H * W;
    float* r = static_cast<float*>(act_buf(d, N, H, W, "r", npix * kOut * 4, rc).p);
    if (rc) return rc;
    bool done = false;
    if ((rc = forward(d, N, H, W, feat, r, s, true, base, out_rgb, &done))) return rc;
    if (done) return RFK_OK;
    NN_HIP(rfk::nn::residual_out(base, 3, r, kOut, npix, out_rgb, s));
    NN_HIP(rfk::nn::clamp_rgb(out_rgb, npix * 3, out_rgb, s));
    return RFK_OK;
}

// The reference's denoise call (interface/denoiser.h:35-38): a tonemapped
// image in, the denoised image out, on the caller's stream. The features come
// from the image itself (the alpha of a 4-channel image carries the density);
// the image is padded to multiples of 8 for the network and cropped back.
int rfk_denoise(rfk_denoiser* d, const void* in, int in_format, void* out, int out_format, uint32_t width,
                uint32_t height, void* stream) {
    if (!d || !in || !out) return fail(RFK_ERR_INVALID, "null argument");
    auto fmt_ok = [](int f) {
        return f == RFK_PIXEL_FLOAT4 || f == RFK_PIXEL_HALF4 || f == RFK_PIXEL_FLOAT3 || f == RFK_PIXEL_HALF3;
    };
    if (!fmt_ok(in_format) || !fmt_ok(out_format))
        return fail(RFK_ERR_INVALID, "formats must be float4, half4, float3 or half3");
    if (width == 0 || height == 0 || width > 32768 || height > 32768) return fail(RFK_ERR_INVALID, "bad image size");
    if (in == out) return fail(RFK_ERR_INVALID, "in and out must not alias");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int W = int(width), H = int(height), Wp = (W + 7) / 8 * 8, Hp = (H + 7) / 8 * 8;
    const size_t npix = size_t(Wp) * Hp;
    int rc = 0;
    void* feat = act_buf(d, 1, Hp, Wp, "dn_feat", npix * kFeat * 2, rc).p;
    float* base = static_cast<float*>(act_buf(d, 1, Hp, Wp, "dn_base", npix * 3 * 4, rc).p);
    float* rgb = static_cast<float*>(act_buf(d, 1, Hp, Wp, "dn_rgb", npix * 3 * 4, rc).p);
    float* r = static_cast<float*>(act_buf(d, 1, Hp, Wp, "r", npix * kOut * 4, rc).p);
    if (rc) return rc;
    // formats share their numbering with the kernels: 0 float4, 1 half4, 2 float3, 3 half3
    NN_HIP(rfk::nn::image_features(in, in_format, W, H, Wp, Hp, kFeat, feat, base, s));
    bool done = false;
    if ((rc = forward(d, 1, Hp, Wp, feat, r, s, true, base, rgb, &done))) return rc;
    if (!done) {
        NN_HIP(rfk::nn::residual_out(base, 3, r, kOut, npix, rgb, s));
        NN_HIP(rfk::nn::clamp_rgb(rgb, npix * 3, rgb, s));
    }
    NN_HIP(rfk::nn::image_output(rgb, Wp, in, in_format, out, out_format, W, H, s));
    return RFK_OK;
}

// One training step on a batch: MSE between base + U-Net(feat) and target
// (FP32 RGB), backward through every layer, Adam update. Returns the loss.
int rfk_denoiser_train_step(rfk_denoiser* d, const void* feat, const float* base, const float* target, int N, int H,
                            int W, float lr, float* loss, void* stream) {
    if (!d || !feat || !base || !target) return fail(RFK_ERR_INVALID, "null argument");
    if (!shape_ok(N, H, W)) return fail(RFK_ERR_INVALID, "height and width must be multiples of 8");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = 0;
    const size_t npix = size_t(N) * H * W;
    float* r = static_cast<float*>(act_buf(d, N, H, W, "r", npix * kOut * 4, rc).p);
    float* out = static_cast<float*>(act_buf(d, N, H, W, "out", npix * 3 * 4, rc).p);
    void* dr = act_buf(d, N, H, W, "dr", npix * kOut * 2, rc).p;
    if (rc) return rc;
    if ((rc = forward(d, N, H, W, feat, r, s))) return rc;
    NN_HIP(rfk::nn::residual_out(base, 3, r, kOut, npix, out, s));
    NN_HIP(hipMemsetAsync(d->loss_acc, 0, 4, s));
    NN_HIP(rfk::nn::mse_grad(out, target, npix, kOut, 1.0f / float(npix * 3), dr, d->loss_acc, s));
    if ((rc = backward(d, N, H, W, feat, dr, s))) return rc;
    d->step++;
    if (lr > 0.f) {
        NN_HIP(rfk::nn::adam(d->w, d->mw, d->vw, d->gw, d->nw, lr, 0.9f, 0.999f, 1e-8f, d->step, 0.f, d->wbf, s));
        NN_HIP(rfk::nn::adam(d->b, d->mb, d->vb, d->gb, d->nb, lr, 0.9f, 0.999f, 1e-8f, d->step, 0.f, nullptr, s));
    }
    float h = 0.f;
    NN_HIP(hipMemcpyAsync(&h, d->loss_acc, 4, hipMemcpyDeviceToHost, s));
    NN_HIP(hipStreamSynchronize(s));
    if (loss) *loss = h / float(npix * 3);
    return RFK_OK;
}

// Single-layer entry points (used by the gradient checks).
int rfk_nn_conv3x3_forward(const void* x1, int c1, const void* x2, int c2, int N, int H, int W, const void* wgt,
                           const float* bias, int Cout, int act, void* y, float* yf, void* stream) {
    return rfk_nn_conv3x3_forward_ex(x1, c1, 0, x2, c2, N, H, W, wgt, bias, Cout, act, y, yf, stream);
}

int rfk_nn_conv3x3_forward_ex(const void* x1, int c1, int x1_mode, const void* x2, int c2, int N, int H, int W,
                              const void* wgt, const float* bias, int Cout, int act, void* y, float* yf, void* stream) {
    if (!x1 || !wgt || (!y && !yf)) return fail(RFK_ERR_INVALID, "null argument");
    if (x1_mode < 0 || x1_mode > 2) return fail(RFK_ERR_INVALID, "x1_mode must be 0, 1 or 2");
    NN_HIP(rfk::nn::conv3x3_forward(x1, c1, x2, c2, N, H, W, wgt, bias, Cout, act, 0.f, y, yf,
                                    static_cast<hipStream_t>(stream), x1_mode));
    return RFK_OK;
}

int rfk_nn_conv3x3_forward_pool(const void* x1, int c1, int x1_mode, const void* x2, int c2, int N, int H, int W,
                                const void* wgt, const float* bias, int Cout, int act, void* y, void* ypool,
                                void* stream) {
    if (!x1 || !wgt || !y || !ypool) return fail(RFK_ERR_INVALID, "null argument");
    if (x1_mode < 0 || x1_mode > 2) return fail(RFK_ERR_INVALID, "x1_mode must be 0, 1 or 2");
    if ((H | W) & 1 || Cout % 8) return fail(RFK_ERR_INVALID, "pooling needs even H and W and Cout a multiple of 8");
    NN_HIP(rfk::nn::conv3x3_forward(x1, c1, x2, c2, N, H, W, wgt, bias, Cout, act, 0.f, y, nullptr,
                                    static_cast<hipStream_t>(stream), x1_mode, ypool));
    return RFK_OK;
}

int rfk_nn_conv3x3_pair_forward(const void* x, int cin, int N, int H, int W, const void* wa, const float* ba,
                                int act_a, const void* wb, const float* bb, int cout_b, int act_b, void* y,
                                void* ypool, const float* base, float* out_rgb, void* stream) {
    const hipError_t e = rfk::nn::conv3x3_pair_forward(x, cin, N, H, W, wa, ba, act_a, wb, bb, cout_b, act_b, y, ypool,
                                                       base, out_rgb, static_cast<hipStream_t>(stream));
    if (e == hipErrorInvalidValue) return fail(RFK_ERR_INVALID, "conv3x3_pair_forward: unsupported shape or arguments");
    if (e != hipSuccess) return fail(RFK_ERR_HIP, std::string("conv3x3_pair_forward: ") + hipGetErrorString(e));
    return RFK_OK;
}

int rfk_nn_avgpool2(const void* x, int N, int H, int W, int C, void* y, void* stream) {
    if (!x || !y) return fail(RFK_ERR_INVALID, "null argument");
    if (N <= 0 || H < 2 || W < 2 || (H | W) & 1 || C <= 0 || C % 8)
        return fail(RFK_ERR_INVALID, "avgpool2 needs even H and W and C a multiple of 8");
    NN_HIP(rfk::nn::avgpool2_forward(x, N, H, W, C, y, static_cast<hipStream_t>(stream)));
    return RFK_OK;
}

int rfk_nn_conv3x3_wgrad(const void* x1, int c1, const void* x2, int c2, const void* dy, int Cout, int N, int H, int W,
                         float* dw, void* stream) {
    size_t ws = rfk::nn::conv3x3_wgrad_workspace(c1, c2, Cout, N, H, W, nullptr, nullptr);
    float* tmp = nullptr;
    NN_HIP(hipMalloc(reinterpret_cast<void**>(&tmp), ws + 256));
    hipError_t e = rfk::nn::conv3x3_wgrad(x1, c1, x2, c2, dy, Cout, N, H, W, dw, 0.f, tmp,
                                          static_cast<hipStream_t>(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(RFK_ERR_HIP, hipGetErrorString(e));
    return RFK_OK;
}

int rfk_nn_conv3x3_dgrad(const void* dy, int Cout, int N, int H, int W, const void* wgt, int Cin, void* dx,
                         void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* wt = nullptr;
    NN_HIP(hipMalloc(&wt, size_t(Cout) * 9 * Cin * 2 + 64));
    hipError_t e = rfk::nn::flip_weights(wgt, Cout, Cin, wt, s);
    if (e == hipSuccess) e = rfk::nn::conv3x3_forward(dy, Cout, nullptr, 0, N, H, W, wt, nullptr, Cin, 0, 0.f, dx,
                                                      nullptr, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(wt);
    if (e != hipSuccess) return fail(RFK_ERR_HIP, hipGetErrorString(e));
    return RFK_OK;
}

int rfk_nn_gather_crops(const void* src, int n_images, int H, int W, int pixel_bytes,
                        const int* offsets /* host, 3 per crop */, int n, int ch, int cw, void* dst, void* stream) {
    if (!src || !offsets || !dst || n <= 0 || pixel_bytes <= 0 || n_images <= 0)
        return fail(RFK_ERR_INVALID, "bad argument");
    for (int k = 0; k < n; k++)
        if (offsets[3 * k] < 0 || offsets[3 * k] >= n_images || offsets[3 * k + 1] < 0 || offsets[3 * k + 2] < 0 || offsets[3 * k + 1] + ch > H || offsets[3 * k + 2] + cw > W)
            return fail(RFK_ERR_INVALID, "crop outside the image");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int* d_off = nullptr;
    NN_HIP(hipMalloc(reinterpret_cast<void**>(&d_off), size_t(n) * 3 * sizeof(int)));
    hipError_t e = hipMemcpyAsync(d_off, offsets, size_t(n) * 3 * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = rfk::nn::gather_crops(src, H, W, pixel_bytes, d_off, n, ch, cw, dst, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_off);
    if (e != hipSuccess) return fail(RFK_ERR_HIP, hipGetErrorString(e));
    return RFK_OK;
}

// Staging for rfk_nn_gather_crops_noalloc: the crop offsets of one call, in
// a file-scope device array (each GPU holds its own copy until the process
// ends). Host code reaches it only through the address hipGetSymbolAddress
// returns for the current device, looked up once per device.
static __device__ __align__(256) int scratch[3 * RFK_NOALLOC_MAX_CROPS];

int rfk_nn_gather_crops_noalloc(const void* src, int n_images, int H, int W, int pixel_bytes, const int* offsets,
                                int n, int ch, int cw, void* dst, void* stream) {
    static std::mutex mu;
    static std::array<void*, 64> dev_ptr{};
    if (!src || !offsets || !dst || n <= 0 || pixel_bytes <= 0 || n_images <= 0)
        return fail(RFK_ERR_INVALID, "bad argument");
    if (n > RFK_NOALLOC_MAX_CROPS)
        return fail(RFK_ERR_INVALID, "more crops than the staging array holds (RFK_NOALLOC_MAX_CROPS)");
    for (int k = 0; k < n; k++)
        if (offsets[3 * k] < 0 || offsets[3 * k] >= n_images || offsets[3 * k + 1] < 0 || offsets[3 * k + 2] < 0 ||
            offsets[3 * k + 1] + ch > H || offsets[3 * k + 2] + cw > W)
            return fail(RFK_ERR_INVALID, "crop outside the image stack");
    int dev = 0;
    NN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= int(dev_ptr.size())) return fail(RFK_ERR_INVALID, "device index out of range");
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!dev_ptr[size_t(dev)]) {
            void* p = nullptr;
            NN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            dev_ptr[size_t(dev)] = p;
        }
        ptr = dev_ptr[size_t(dev)];
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    NN_HIP(hipMemcpyAsync(ptr, offsets, size_t(n) * 3 * sizeof(int), hipMemcpyHostToDevice, s));
    NN_HIP(rfk::nn::gather_crops(src, H, W, pixel_bytes, static_cast<const int*>(ptr), n, ch, cw, dst, s));
    NN_HIP(hipStreamSynchronize(s));
    return RFK_OK;
}

int rfk_nn_inject_alloc_failures(int count) {
    g_inject_alloc_failures.store(count < 0 ? 0 : count);
    return RFK_OK;
}

}  // extern "C"
