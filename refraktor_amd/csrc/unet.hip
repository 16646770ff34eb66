// Denoiser kernels: 3x3 "same" convolutions on BF16 NHWC activations with
// FP32 accumulation on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16), and
// the small streaming kernels a U-Net needs (pooling, upsampling, activation
// gradients, loss, Adam, feature preparation). The librefrakt denoisers are
// third-party binaries (OIDN / OptiX); this network has no reference
// counterpart and is checked against the oracle's float64 convolution.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <unordered_map>

#include "unet_kernels.hpp"

namespace rfk::nn {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int KC = 32;       // input channels staged per step
constexpr int PAD = 8;       // LDS row padding (bf16) against bank conflicts

__device__ __forceinline__ float act_fwd(float v, int act, float slope) {
    if (act == 1) return v > 0.f ? v : 0.f;
    if (act == 2) return v > 0.f ? v : v * slope;
    return v;
}

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = bf16(0.0f);
    return z;
}

// Raw buffer reads: a lane whose byte offset lies past the resource's size
// reads zeros, so a masked-off lane passes kBufferOOB (beyond any resource,
// all of which hold fewer than 2^31 bytes) instead of selecting zero after
// the load, and the load stays in flight until its data is used.
constexpr unsigned kBufferOOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, int(bytes), 0x00020000);
}
__device__ __forceinline__ bf16x8 buffer_load8(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0));
}
// a store at kBufferOOB (or into a resource of size 0) writes nothing
__device__ __forceinline__ void buffer_store8(bf16x8 v, __amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, byte_off, 0, 0);
}

// act_fwd with the activation's branch taken out of the lanes: relu (act 1)
// gives 0 below zero, leaky (2) v * slope, none (0) v * 1
__device__ __forceinline__ float act_select(float v, bool relu, float neg_slope) {
    return v > 0.f ? v : (relu ? 0.f : v * neg_slope);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// one lane's 16 accumulator values (channels 8i + 4kh + j) plus their
// biases (bias + 8i + 4kh + j), activated and rounded to BF16; the ReLU of
// the hidden layers takes one max per value, the other activations go
// through act_select
__device__ __forceinline__ void act16_bf16(const f32x16& acc, const float* bias, int kh, bool relu, float neg_slope,
                                           bf16x4 out[4]) {
    if (relu) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float4 b = *reinterpret_cast<const float4*>(bias + 8 * i + 4 * kh);
            out[i][0] = bf16(fmaxf(acc[i * 4 + 0] + b.x, 0.f));
            out[i][1] = bf16(fmaxf(acc[i * 4 + 1] + b.y, 0.f));
            out[i][2] = bf16(fmaxf(acc[i * 4 + 2] + b.z, 0.f));
            out[i][3] = bf16(fmaxf(acc[i * 4 + 3] + b.w, 0.f));
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float4 b = *reinterpret_cast<const float4*>(bias + 8 * i + 4 * kh);
            out[i][0] = bf16(act_select(acc[i * 4 + 0] + b.x, false, neg_slope));
            out[i][1] = bf16(act_select(acc[i * 4 + 1] + b.y, false, neg_slope));
            out[i][2] = bf16(act_select(acc[i * 4 + 2] + b.z, false, neg_slope));
            out[i][3] = bf16(act_select(acc[i * 4 + 3] + b.w, false, neg_slope));
        }
    }
}

// ------------------------------------------------------------- forward
// Output tile of a workgroup: 128*R consecutive pixels of one row x 32*C
// output channels; wave w computes pixels [32R*w, +32R) x all 32C channels as
// R x C 32x32 MFMA accumulators (A = pixels x K from the input halo, B =
// K x channels from the weights, K = KCH input channels of one tap per step).
// Input = virtual concat of x1 (c1 channels) and x2 (c2). x1 may be read
// through a 2x upsample (x1 stored at H/2 x W/2, nearest) or a 2x2 average
// pool (x1 stored at 2H x 2W), so the decoder and encoder skip the
// intermediate tensors. It serves the layers the strip kernel below does not
// take (inputs too wide for the strip kernel's weights in LDS, x1 ending
// inside a 32-channel chunk, tensors of 2^30 bytes or more). Shapes: (R, C, KCH) = (1,1,32) and (1,1,16) for
// Cout <= 32, (1,2,32) for <= 64, (1,3,16) for <= 96, (1,4,16) above.
__device__ __forceinline__ bf16x8 pool8(const bf16* p00, const bf16* p01, const bf16* p10, const bf16* p11) {
    const bf16x8 a = load8(p00), b = load8(p01), c = load8(p10), d = load8(p11);
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = bf16(0.25f * (float(a[i]) + float(b[i]) + float(c[i]) + float(d[i])));
    return o;
}

template <int R, int C, int KCH>
__global__ __launch_bounds__(256) void conv3x3_fwd_kernel(const bf16* __restrict__ x1, int c1, int mode1,
                                                          const bf16* __restrict__ x2, int c2, int N, int H, int W,
                                                          const bf16* __restrict__ wgt, const float* __restrict__ bias,
                                                          int Cout, int act, float slope, bf16* __restrict__ y,
                                                          float* __restrict__ yf) {
    constexpr int BPX = 128 * R, TC = 32 * C, KP = KCH + PAD, G = KCH / 8;
    __shared__ __attribute__((aligned(16))) bf16 s_in[3][BPX + 2][KP];
    __shared__ __attribute__((aligned(16))) bf16 s_w[9][TC][KP];
    const int Cin = c1 + c2;
    const int x0 = int(blockIdx.x) * BPX;
    const int yy = int(blockIdx.y);
    const int ntile = (Cout + TC - 1) / TC;
    const int n = int(blockIdx.z) / ntile;
    const int co0 = (int(blockIdx.z) % ntile) * TC;
    const int t = int(threadIdx.x);
    const int wv = t >> 6, l = t & 63;
    const int m = l & 31, kh = l >> 5;
    const int H1 = mode1 == 1 ? H / 2 : (mode1 == 2 ? 2 * H : H), W1 = mode1 == 1 ? W / 2 : (mode1 == 2 ? 2 * W : W);

    f32x16 acc[R][C];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[r][c][i] = 0.f;

    // staging: every thread first issues all of its global loads, then writes
    // them to LDS, so the loads of a chunk are in flight together
    constexpr int NIN = (3 * (BPX + 2) * G + 255) / 256, NW = (9 * TC * G + 255) / 256;
    for (int c0 = 0; c0 < Cin; c0 += KCH) {
        bf16x8 vin[NIN], vw[NW];
#pragma unroll
        for (int q = 0; q < NIN; q++) {
            const int v = t + q * 256;
            const int g = v % G, col = (v / G) % (BPX + 2), rr = v / (G * (BPX + 2));
            const int gy = yy + rr - 1, gx = x0 + col - 1, c = c0 + g * 8;
            bf16x8 val = zero8();
            if (rr < 3 && gy >= 0 && gy < H && gx >= 0 && gx < W && c < Cin) {
                if (c < c1) {
                    if (mode1 == 1) {
                        val = load8(x1 + ((size_t(n) * H1 + gy / 2) * W1 + gx / 2) * c1 + c);
                    } else if (mode1 == 2) {
                        const size_t r0 = (size_t(n) * H1 + 2 * gy) * W1 + 2 * gx, r1 = r0 + W1;
                        val = pool8(x1 + r0 * c1 + c, x1 + (r0 + 1) * c1 + c, x1 + r1 * c1 + c, x1 + (r1 + 1) * c1 + c);
                    } else {
                        val = load8(x1 + ((size_t(n) * H + gy) * W + gx) * c1 + c);
                    }
                } else {
                    val = load8(x2 + ((size_t(n) * H + gy) * W + gx) * c2 + (c - c1));
                }
            }
            vin[q] = val;
        }
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int v = t + q * 256;
            const int g = v % G, co = (v / G) % TC, tap = v / (G * TC);
            const int c = c0 + g * 8, o = co0 + co;
            bf16x8 val = zero8();
            if (tap < 9 && o < Cout && c < Cin) val = load8(wgt + (size_t(o) * 9 + tap) * Cin + c);
            vw[q] = val;
        }
#pragma unroll
        for (int q = 0; q < NIN; q++) {
            const int v = t + q * 256;
            const int g = v % G, col = (v / G) % (BPX + 2), rr = v / (G * (BPX + 2));
            if (rr < 3) *reinterpret_cast<bf16x8*>(&s_in[rr][col][g * 8]) = vin[q];
        }
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int v = t + q * 256;
            const int g = v % G, co = (v / G) % TC, tap = v / (G * TC);
            if (tap < 9) *reinterpret_cast<bf16x8*>(&s_w[tap][co][g * 8]) = vw[q];
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int ks = 0; ks < KCH / 16; ks++) {
                bf16x8 b[C];
#pragma unroll
                for (int c = 0; c < C; c++) b[c] = *reinterpret_cast<const bf16x8*>(&s_w[tap][c * 32 + m][ks * 16 + kh * 8]);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const bf16x8 av =
                        *reinterpret_cast<const bf16x8*>(&s_in[dy][wv * 32 * R + r * 32 + m + dx][ks * 16 + kh * 8]);
#pragma unroll
                    for (int c = 0; c < C; c++) acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b[c], acc[r][c], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D[row = pixel][col = channel]: lane holds col m, rows 8i + 4kh + j
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int co = co0 + c * 32 + m;
        if (co >= Cout) continue;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int px = x0 + wv * 32 * R + r * 32 + i * 8 + kh * 4 + j;
                    if (px >= W) continue;
                    const float v = act_fwd(acc[r][c][i * 4 + j] + bv, act, slope);
                    const size_t o = ((size_t(n) * H + yy) * W + px) * Cout + co;
                    if (y) y[o] = bf16(v);
                    if (yf) yf[o] = v;
                }
            }
        }
    }
}

// --------------------------------------------------------- weight grad
// dW[co][tap][ci] = sum_p dY[p][co] * X[p + off(tap)][ci], split over pixel
// groups (blockIdx.x) and written as partial sums; one WG owns a 32x32
// (co, ci) tile and all 9 taps (wave w: taps w, w+4, w+8).
constexpr int WG_PX = 64;

__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(const bf16* __restrict__ x1, int c1,
                                                            const bf16* __restrict__ x2, int c2,
                                                            const bf16* __restrict__ dy, int Cout, int N, int H,
                                                            int W, int tiles_per_group, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) bf16 s_dy[32][WG_PX + PAD];
    __shared__ __attribute__((aligned(16))) bf16 s_x[3][3][32][WG_PX + PAD];
    const int Cin = c1 + c2;
    const int co0 = int(blockIdx.y) * 32, ci0 = int(blockIdx.z) * 32;
    const int t = int(threadIdx.x), wv = t >> 6, l = t & 63, m = l & 31, kh = l >> 5;
    const int tiles_x = (W + WG_PX - 1) / WG_PX;
    const long total_tiles = long(N) * H * tiles_x;

    f32x16 acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[q][i] = 0.f;

    for (int it = 0; it < tiles_per_group; it++) {
        const long tile = long(blockIdx.x) * tiles_per_group + it;
        if (tile >= total_tiles) break;
        const int tx = int(tile % tiles_x);
        const int row = int((tile / tiles_x) % H);
        const int n = int(tile / (long(tiles_x) * H));
        const int x0 = tx * WG_PX;
        // dY^T tile: s_dy[co][px]
        for (int v = t; v < WG_PX * 4; v += 256) {
            const int g = v & 3, px = v >> 2, co = co0 + g * 8;
            bf16x8 val = zero8();
            if (x0 + px < W && co < Cout) val = load8(dy + ((size_t(n) * H + row) * W + x0 + px) * Cout + co);
#pragma unroll
            for (int e = 0; e < 8; e++) s_dy[g * 8 + e][px] = val[e];
        }
        // X^T tiles shifted by dx: s_x[dy][dx][ci][p] = X[row+dy-1][x0+p+dx-1][ci]
        for (int v = t; v < 3 * (WG_PX + 2) * 4; v += 256) {
            const int g = v & 3, col = (v >> 2) % (WG_PX + 2), r = (v >> 2) / (WG_PX + 2);
            const int gy = row + r - 1, gx = x0 + col - 1, c = ci0 + g * 8;
            bf16x8 val = zero8();
            if (gy >= 0 && gy < H && gx >= 0 && gx < W && c < Cin) {
                const size_t pix = (size_t(n) * H + gy) * W + gx;
                val = c < c1 ? load8(x1 + pix * c1 + c) : load8(x2 + pix * c2 + (c - c1));
            }
#pragma unroll
            for (int ddx = 0; ddx < 3; ddx++) {
                const int p = col - ddx;
                if (p < 0 || p >= WG_PX) continue;
#pragma unroll
                for (int e = 0; e < 8; e++) s_x[r][ddx][g * 8 + e][p] = val[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int tap = wv + 4 * q;
            if (tap < 9) {
                const int ddy = tap / 3, ddx = tap % 3;
#pragma unroll
                for (int ks = 0; ks < WG_PX / 16; ks++) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&s_dy[m][ks * 16 + kh * 8]);
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&s_x[ddy][ddx][m][ks * 16 + kh * 8]);
                    acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // D[row = co][col = ci]
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int tap = wv + 4 * q;
        if (tap >= 9) continue;
        const int ci = ci0 + m;
        if (ci >= Cin) continue;
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int co = co0 + i * 8 + kh * 4 + j;
                if (co >= Cout) continue;
                partial[((size_t(blockIdx.x) * Cout + co) * 9 + tap) * Cin + ci] = acc[q][i * 4 + j];
            }
    }
}

__global__ void reduce_partials_kernel(const float* __restrict__ partial, int groups, size_t count,
                                       float* __restrict__ out, float beta) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    float s = 0.f;
    for (int k = 0; k < groups; k++) s += partial[size_t(k) * count + g];
    // beta 0 overwrites: the output may hold anything, NaN included
    out[g] = beta == 0.f ? s : beta * out[g] + s;
}

// dbias[co] = sum over pixels of dY; one block per channel group of 64
__global__ void bias_grad_kernel(const bf16* __restrict__ dy, size_t npix, int C, float* __restrict__ db, float beta) {
    __shared__ float s[256];
    const int c = int(blockIdx.x);
    float acc = 0.f;
    for (size_t p = threadIdx.x; p < npix; p += blockDim.x) acc += float(dy[p * C + c]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = beta == 0.f ? s[0] : beta * db[c] + s[0];
}

// W'[ci][tap][co] = W[co][8 - tap][ci]: the data gradient is the forward
// convolution of dY with these weights
__global__ void flip_weights_kernel(const bf16* __restrict__ w, int Cout, int Cin, bf16* __restrict__ wt) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= size_t(Cout) * 9 * Cin) return;
    const int ci = int(g % Cin), tap = int((g / Cin) % 9), co = int(g / (size_t(Cin) * 9));
    wt[(size_t(ci) * 9 + (8 - tap)) * Cout + co] = w[g];
}

// --------------------------------------------------------- elementwise
__global__ void act_bwd_kernel(const bf16* __restrict__ y, const bf16* __restrict__ dy, size_t n, int act, float slope,
                               bf16* __restrict__ dx) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float yv = float(y[g]), d = float(dy[g]);
    float r = d;
    if (act == 1) r = yv > 0.f ? d : 0.f;
    else if (act == 2) r = yv > 0.f ? d : d * slope;
    dx[g] = bf16(r);
}

// 8 channels of one output pixel per thread (C is a multiple of 8)
__global__ void avgpool2_fwd_kernel(const bf16* __restrict__ x, int N, int H, int W, int C, bf16* __restrict__ y) {
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= size_t(N) * Ho * Wo * C8) return;
    const int c = int(g % C8) * 8, xo = int((g / C8) % Wo), yo = int((g / (size_t(C8) * Wo)) % Ho),
              n = int(g / (size_t(C8) * Wo * Ho));
    const size_t r0 = (size_t(n) * H + 2 * yo) * W + 2 * xo, r1 = r0 + W;
    *reinterpret_cast<bf16x8*>(y + g * 8) = pool8(x + r0 * C + c, x + (r0 + 1) * C + c, x + r1 * C + c, x + (r1 + 1) * C + c);
}

__global__ void avgpool2_bwd_kernel(const bf16* __restrict__ dy, int N, int H, int W, int C, bf16* __restrict__ dx) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= size_t(N) * H * W * C) return;
    const int c = int(g % C), xx = int((g / C) % W), yy = int((g / (size_t(C) * W)) % H), n = int(g / (size_t(C) * W * H));
    const int Ho = H / 2, Wo = W / 2;
    const int yo = yy / 2, xo = xx / 2;
    float v = 0.f;
    if (yo < Ho && xo < Wo) v = 0.25f * float(dy[((size_t(n) * Ho + yo) * Wo + xo) * C + c]);
    dx[g] = bf16(v);
}

__global__ void upsample2_fwd_kernel(const bf16* __restrict__ x, int N, int H, int W, int C, bf16* __restrict__ y) {
    const int Ho = H * 2, Wo = W * 2;
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= size_t(N) * Ho * Wo * C) return;
    const int c = int(g % C), xo = int((g / C) % Wo), yo = int((g / (size_t(C) * Wo)) % Ho),
              n = int(g / (size_t(C) * Wo * Ho));
    y[g] = x[((size_t(n) * H + yo / 2) * W + xo / 2) * C + c];
}

__global__ void upsample2_bwd_kernel(const bf16* __restrict__ dy, int N, int H, int W, int C, bf16* __restrict__ dx) {
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= size_t(N) * H * W * C) return;
    const int c = int(g % C), xx = int((g / C) % W), yy = int((g / (size_t(C) * W)) % H), n = int(g / (size_t(C) * W * H));
    const int Ho = H * 2, Wo = W * 2;
    auto at = [&](int a, int b) { return float(dy[((size_t(n) * Ho + a) * Wo + b) * C + c]); };
    dx[g] = bf16(at(2 * yy, 2 * xx) + at(2 * yy, 2 * xx + 1) + at(2 * yy + 1, 2 * xx) + at(2 * yy + 1, 2 * xx + 1));
}

// split a concatenated gradient [.., c1 + c2] into dx1 = g[..:c1] and
// dx2 += g[c1:] (the skip connection also receives the encoder's gradient)
__global__ void split_add_kernel(const bf16* __restrict__ g, size_t npix, int c1, int c2, bf16* __restrict__ dx1,
                                 bf16* __restrict__ dx2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int C = c1 + c2;
    if (i >= npix * C) return;
    const size_t p = i / C;
    const int c = int(i % C);
    if (c < c1) {
        if (dx1) dx1[p * c1 + c] = g[i];
    } else if (dx2) {
        const size_t o = p * c2 + (c - c1);
        dx2[o] = bf16(float(dx2[o]) + float(g[i]));
    }
}

__global__ void add_bf16_kernel(const bf16* __restrict__ a, bf16* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf16(float(a[i]) + float(b[i]));
}

// out (N,H,W,3) = base RGB + residual (first 3 channels of r, Cr stride);
// loss = mean squared error vs target; gradient of the loss w.r.t. r
__global__ void residual_out_kernel(const float* __restrict__ base, int Cb, const float* __restrict__ r, int Cr,
                                    size_t npix, float* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npix * 3) return;
    const size_t p = i / 3;
    const int c = int(i % 3);
    out[i] = base[p * Cb + c] + r[p * Cr + c];
}

__global__ void mse_grad_kernel(const float* __restrict__ out, const float* __restrict__ target, size_t npix, int Cr,
                                float scale, bf16* __restrict__ dr, float* __restrict__ loss_acc) {
    __shared__ float s[256];
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < npix * size_t(Cr)) {
        const size_t p = i / Cr;
        const int c = int(i % Cr);
        float g = 0.f;
        if (c < 3) {
            const float d = out[p * 3 + c] - target[p * 3 + c];
            e2 = d * d;
            g = 2.f * d * scale;
        }
        dr[i] = bf16(g);
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_acc, s[0]);
}

// Adam on FP32 master weights, refreshing the BF16 copy used by the kernels
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, size_t n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2, float wd, bf16* __restrict__ wb) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (wb) wb[i] = bf16(wi);
}

__global__ void f32_to_bf16_kernel(const float* __restrict__ a, bf16* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf16(a[i]);
}

__global__ void bf16_to_f32_kernel(const bf16* __restrict__ a, float* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

// Denoiser input features from a fixed-point histogram and its tonemapped
// image: [R, G, B, alpha] of the tonemapped frame, log2(1 + density) / 16,
// log2(1 + hits) / 16 and zero padding up to Cpad channels; also the FP32
// base image the network adds its residual to.
// `hist` holds u32 RGBA bins, or u64 RGBA bins when hist64 is set
__global__ void features_kernel(const uint32_t* __restrict__ hist, bool hist64, const float4* __restrict__ tm,
                                size_t npix, int Cpad, bf16* __restrict__ feat, float* __restrict__ base) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float w = hist64 ? float(reinterpret_cast<const unsigned long long*>(hist)[4 * p + 3])
                           : float(hist[4 * p + 3]);
    const float4 c = tm[p];
    const float dens = w * (1.0f / 255.0f);
    float f[8] = {c.x, c.y, c.z, c.w, log2f(1.0f + dens) * (1.0f / 16.0f), w > 0.0f ? 1.0f : 0.0f, 0.f, 0.f};
    for (int k = 0; k < Cpad; k++) feat[p * Cpad + k] = bf16(k < 6 ? f[k] : 0.f);
    base[p * 3 + 0] = c.x;
    base[p * 3 + 1] = c.y;
    base[p * 3 + 2] = c.z;
}

// Denoiser input from a tonemapped image alone (the reference's denoise call
// takes images): RGB as they are, the density from the tonemap's alpha,
// sat(log2(w) / 24) (tonemap.cu:95-98), as w = 2^(24 alpha); a covered pixel
// with alpha 0 holds at most one draw and counts as one, and a 3-channel
// image, which has no alpha, gives every covered pixel one draw. The image is
// padded with empty pixels to Wp x Hp (multiples of 8). fmt: 0 float4,
// 1 half4, 2 float3, 3 half3.
__device__ __forceinline__ float4 load_pixel(const void* img, int fmt, size_t i) {
    switch (fmt) {
        case 0: return reinterpret_cast<const float4*>(img)[i];
        case 1: {
            const __half* h = reinterpret_cast<const __half*>(img) + 4 * i;
            return {__half2float(h[0]), __half2float(h[1]), __half2float(h[2]), __half2float(h[3])};
        }
        case 2: {
            const float* f = reinterpret_cast<const float*>(img) + 3 * i;
            return {f[0], f[1], f[2], 0.f};
        }
        default: {
            const __half* h = reinterpret_cast<const __half*>(img) + 3 * i;
            return {__half2float(h[0]), __half2float(h[1]), __half2float(h[2]), 0.f};
        }
    }
}

__global__ void image_features_kernel(const void* __restrict__ img, int fmt, int W, int H, int Wp, int Hp, int Cpad,
                                      bf16* __restrict__ feat, float* __restrict__ base) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= size_t(Wp) * Hp) return;
    const int y = int(p / size_t(Wp)), x = int(p % size_t(Wp));
    float4 c = {0.f, 0.f, 0.f, 0.f};
    if (x < W && y < H) c = load_pixel(img, fmt, size_t(y) * W + x);
    const bool covered = c.w > 0.f || c.x > 0.f || c.y > 0.f || c.z > 0.f;
    const float dens = c.w > 0.f ? exp2f(24.0f * c.w) : (covered ? 1.0f : 0.0f);
    const float f[8] = {c.x, c.y, c.z, c.w, log2f(1.0f + dens) * (1.0f / 16.0f), covered ? 1.0f : 0.0f, 0.f, 0.f};
    for (int k = 0; k < Cpad; k++) feat[p * Cpad + k] = bf16(k < 6 ? f[k] : 0.f);
    base[p * 3 + 0] = c.x;
    base[p * 3 + 1] = c.y;
    base[p * 3 + 2] = c.z;
}

// The denoised RGB (Wp x Hp, FP32) into the caller's image format; the alpha
// of a 4-channel output is the input's (1 when the input has none).
__global__ void image_output_kernel(const float* __restrict__ rgb, int Wp, const void* __restrict__ img, int in_fmt,
                                    void* __restrict__ out, int out_fmt, int W, int H) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(W) * H) return;
    const int y = int(i / size_t(W)), x = int(i % size_t(W));
    const float* c = rgb + (size_t(y) * Wp + x) * 3;
    const float a = (in_fmt == 0 || in_fmt == 1) ? load_pixel(img, in_fmt, i).w : 1.0f;
    switch (out_fmt) {
        case 0: reinterpret_cast<float4*>(out)[i] = {c[0], c[1], c[2], a}; break;
        case 1: {
            __half* h = reinterpret_cast<__half*>(out) + 4 * i;
            h[0] = __float2half(c[0]); h[1] = __float2half(c[1]); h[2] = __float2half(c[2]); h[3] = __float2half(a);
        } break;
        case 2: {
            float* f = reinterpret_cast<float*>(out) + 3 * i;
            f[0] = c[0]; f[1] = c[1]; f[2] = c[2];
        } break;
        default: {
            __half* h = reinterpret_cast<__half*>(out) + 3 * i;
            h[0] = __float2half(c[0]); h[1] = __float2half(c[1]); h[2] = __float2half(c[2]);
        } break;
    }
}

__global__ void clamp_rgb_kernel(const float* __restrict__ in, size_t n, float* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fminf(fmaxf(in[i], 0.f), 1.f);
}

// copy n crops of ch x cw pixels (C channels of elem_bytes each) out of a
// stack of H x W images; crop k starts at (img[k], y[k], x[k]) = offs[3k..]
__global__ void gather_crops_kernel(const unsigned char* __restrict__ src, int H, int W, int row_bytes,
                                    const int* __restrict__ offs, int ch, int cw, unsigned char* __restrict__ dst) {
    const int k = int(blockIdx.z), r = int(blockIdx.y);
    const int img = offs[3 * k], y0 = offs[3 * k + 1], x0 = offs[3 * k + 2];
    const size_t srow = (size_t(img) * H + y0 + r) * W + x0;
    const size_t drow = (size_t(k) * ch + r) * cw;
    for (int b = int(blockIdx.x) * blockDim.x + threadIdx.x; b < cw * row_bytes; b += gridDim.x * blockDim.x)
        dst[drow * row_bytes + b] = src[srow * row_bytes + b];
}

unsigned nb(size_t n, unsigned b = 256) { return unsigned((n + b - 1) / b); }

}  // namespace

// Eight channels [c, c + 8) of input pixel (gy, gx) of the virtual concat
// (x1 through MODE: 0 direct, 1 nearest 2x upsample, 2 2x2 average pool;
// then x2), zero where `ok` is false. Branch-free: the coordinates and the
// channel are clamped into the tensors first, so the load is always in
// bounds and a select applies `ok` afterwards. 32-bit offsets inside one
// image (x1n, x2n point at the image).
template <int MODE>
__device__ __forceinline__ bf16x8 load_px8(const bf16* __restrict__ x1n, int c1, const bf16* __restrict__ x2n, int c2,
                                           int H, int W, int gy, int gx, int c, bool ok) {
    gy = min(max(gy, 0), H - 1);
    gx = min(max(gx, 0), W - 1);
    const bool use2 = c >= c1 && c2 > 0;
    bf16x8 v;
    if (MODE == 0) {
        const int cx = use2 ? min(c - c1, c2 - 8) : min(c, c1 - 8);
        const bf16* p = use2 ? x2n + unsigned((gy * W + gx) * c2 + cx) : x1n + unsigned((gy * W + gx) * c1 + cx);
        v = load8(p);
    } else if (MODE == 1) {
        const int cx = use2 ? min(c - c1, c2 - 8) : min(c, c1 - 8);
        const bf16* p = use2 ? x2n + unsigned((gy * W + gx) * c2 + cx)
                             : x1n + unsigned(((gy >> 1) * (W >> 1) + (gx >> 1)) * c1 + cx);
        v = load8(p);
    } else {
        if (use2) {
            v = load8(x2n + unsigned((gy * W + gx) * c2 + min(c - c1, c2 - 8)));
        } else {
            const int cx = min(c, c1 - 8), W1 = 2 * W;
            const unsigned r0 = unsigned(((2 * gy) * W1 + 2 * gx) * c1 + cx), r1 = r0 + unsigned(W1 * c1);
            v = pool8(x1n + r0, x1n + r0 + c1, x1n + r1, x1n + r1 + c1);
        }
    }
    return ok ? v : zero8();
}

// bf16 elements of one strip wave's LDS: the larger of a staged chunk (34
// pixels x KC channels) and an output row (32 pixels x 32 channels), both
// with rows padded by PAD
constexpr int strip_wave_lds(int kc) { return 34 * (kc + PAD) > 32 * (32 + PAD) ? 34 * (kc + PAD) : 32 * (32 + PAD); }

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// COUNT 16-byte fragments into LDS by NT threads: thread t takes fragments
// t, t + NT, ... = src(e), written by dst(e, value). All of a thread's loads
// are in flight before its first LDS write, so staging a layer's weights
// costs one memory latency instead of one per fragment.
template <int NT, int COUNT, class Src, class Dst>
__device__ __forceinline__ void stage_fragments(int t, Src&& src, Dst&& dst) {
    constexpr int PER = (COUNT + NT - 1) / NT;
    bf16x8 v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int e = t + k * NT;
        v[k] = (COUNT % NT == 0 || e < COUNT) ? src(e) : zero8();
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int e = t + k * NT;
        if (COUNT % NT == 0 || e < COUNT) dst(e, v[k]);
    }
}

// The column of a staged pixel row that lane l fills in round q, G lanes per
// pixel (one 16-byte channel group each) and 64 / G columns per round. LDS
// serves a wave's 16-byte accesses sixteen lanes at a time; through rows of
// 40 BF16 (G = 4) lanes 0-15 in column order would take columns 0-3, whose
// bank ranges overlap, so consecutive lane quads take columns four apart
// (0, 4, 8, 12 | 1, 5, 9, 13 | ...) and every group covers all banks once;
// with G = 2 (rows of 24 BF16) the groups take every other column. Measured
// on one MI355X, 16-byte writes in this order run at the rate of linear ones,
// 1.2x faster than in column order, and reads 1.4x faster.
template <int G>
__device__ __forceinline__ int stage_col(int q, int l) {
    const int k = l / G;
    if constexpr (G == 4) {
        return 16 * q + 4 * (k & 3) + (k >> 2);
    } else if constexpr (G == 2) {
        const int j = k >> 3;
        return 32 * q + (j >> 1) * 16 + (j & 1) + 2 * (k & 7);
    } else {
        return (l + 64 * q) / G;
    }
}

// Every wave walks its own 32-pixel column strip down `rows` output rows for
// 32 output channels (grid.z = images x 32-channel tiles), input-stationary:
// each staged input row (one KC-channel chunk at a time, NCH chunks) feeds
// the three output rows it touches, held in three rotating accumulators, so
// an input row is read from LDS once per (dx, k-step) instead of three
// times, and waves never wait for each other after the weights are staged.
// D = W (32 out channels x K) * X (K x 32 pixels): a lane ends up with 16
// channels of one pixel.
// Each chunk's pixels are fetched three chunks ahead with raw buffer loads
// into one of three register sets: a lane outside the image or past the
// channels passes an offset beyond the buffer and reads zeros, so nothing
// waits on the data until it is written to LDS; a lane's column and channel
// group are fixed, so its byte offsets are computed once and each chunk only
// adds the row's and chunk's scalar part. Inside a chunk the MFMA operands
// are read from LDS one (dx, k-step) ahead of their use. Whole 32-channel
// BF16 output tiles leave through the wave's LDS rows as 16-byte stores.
// MODE is x1's read (0 direct, 1 2x upsample, 2 2x2 average pool); byte
// offsets inside one image are 32-bit (the host checks every image's tensors
// hold fewer than 2^30 bytes). NW waves per workgroup share the layer's
// weights in LDS.
// OUT is the store: 0 BF16 rows of whole 32-channel tiles through LDS, 1 the
// same plus their 2x2 average pool, 2 anything else (FP32 output, the rgb
// output layer, channel counts that end inside a tile). With OUT 0 and 1
// every wave issues the same global loads and stores in the same order
// whatever its position: rows past the strip or the image and columns past
// the image load and store through buffer offsets beyond the resource (reads
// give zeros, writes are dropped), so the compiler's wait counts before each
// staged chunk only wait for that chunk's loads, not for the ones fetched
// after it.
template <int KC, int NCH, int MODE, int NW, int OUT>
__global__ __launch_bounds__(64 * NW) void conv3x3_strip_kernel(const bf16* __restrict__ x1, int c1,
                                                            const bf16* __restrict__ x2, int c2, int H, int W,
                                                            const bf16* __restrict__ wgt,
                                                            const float* __restrict__ bias, int Cout, int act,
                                                            float slope, int rows, bf16* __restrict__ y,
                                                            bf16* __restrict__ ypool, float* __restrict__ yf,
                                                            const float* __restrict__ rbase,
                                                            float* __restrict__ rgb) {
    constexpr int SPX = 34, XP = KC + PAD, G = KC / 8, NQ = (SPX * G + 63) / 64, KS = KC / 16, NS = 3 * KS;
    constexpr int WP = NCH * KC + PAD, GW = NCH * KC / 8;
    // this wave's LDS: a staged chunk [SPX][XP], or an output row [32][OXP]
    constexpr int OXP = 32 + PAD, WAVE_LDS = strip_wave_lds(KC);
    static_assert(64 % G == 0, "a lane's channel group is the same in all its loads");
    extern __shared__ __attribute__((aligned(16))) unsigned char conv_smem[];
    const int Cin = c1 + c2;
    const int t = int(threadIdx.x), wv = t >> 6, l = t & 63, m = l & 31, kh = l >> 5;
    // [32 out][9 taps][WP]: every tap and chunk of a lane's output channel
    // lies within an immediate offset of one base address
    bf16* s_w = reinterpret_cast<bf16*>(conv_smem);
    float* s_bias = reinterpret_cast<float*>(s_w + 9 * 32 * WP);    // the tile's 32 biases
    bf16* s_x = s_w + 9 * 32 * WP + 64 + wv * WAVE_LDS;
    const int x0 = int(blockIdx.x) * 32 * NW + wv * 32;  // the wave's first output pixel
    const int y0 = int(blockIdx.y) * rows;
    const int ntile = (Cout + 31) / 32;
    const int n = int(blockIdx.z) / ntile, co0 = (int(blockIdx.z) % ntile) * 32;
    const int H1 = MODE == 1 ? H / 2 : (MODE == 2 ? 2 * H : H), W1 = MODE == 1 ? W / 2 : (MODE == 2 ? 2 * W : W);
    const bf16* x1n = x1 + size_t(n) * H1 * W1 * c1;
    const bf16* x2n = x2 ? x2 + size_t(n) * H * W * c2 : nullptr;
    const __amdgpu_buffer_rsrc_t r1 = buffer_rsrc(x1n, unsigned(H1 * W1 * c1) * 2u);
    const __amdgpu_buffer_rsrc_t r2 = buffer_rsrc(x2n ? x2n : x1n, x2n ? unsigned(H * W * c2) * 2u : 0u);

    // the lane's part of a staged row: channel group g, column stage_col(q, l) in round q
    const int g = l % G;
    unsigned o1c[NQ], o2c[NQ];
    bool colok[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int col = stage_col<G>(q, l), gx = x0 - 1 + col;
        colok[q] = col < SPX && gx >= 0 && gx < W;
        o1c[q] = unsigned((MODE == 1 ? (gx >> 1) : gx) * c1 + g * 8) * 2u;
        o2c[q] = unsigned(gx * c2 + g * 8 - c1) * 2u;  // x2's channel is c - c1
    }
    // (input row jj of the strip = image row y0 - 1 + jj, chunk CH) -> registers;
    // zeros where `live` is false
    auto load_chunk = [&](int jj, auto chc, bf16x8* v, bool live) __attribute__((always_inline)) {
        constexpr int CH = decltype(chc)::value, CB = CH * KC;
        const int gy = y0 - 1 + jj;
        const bool row_ok = live && gy >= 0 && gy < H;
        const bool ch_ok = g * 8 < Cin - CB;
        if (MODE == 2) {
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int col = stage_col<G>(q, l);
                v[q] = load_px8<2>(x1n, c1, x2n, c2, H, W, gy, x0 - 1 + col, CB + g * 8, row_ok && colok[q] && ch_ok);
            }
            return;
        }
        // every chunk lies wholly inside x1 or x2 (the host sends layers
        // with c2 > 0 here only when c1 is a multiple of KC): one resource,
        // chosen per wave
        const bool in2 = c2 > 0 && CB >= c1;
        const unsigned srow = in2 ? unsigned(gy * W * c2 + CB) * 2u
                                  : unsigned((MODE == 1 ? (gy >> 1) * W1 : gy * W) * c1 + CB) * 2u;
        const __amdgpu_buffer_rsrc_t r = in2 ? r2 : r1;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const bool ok = row_ok && colok[q] && ch_ok;
            v[q] = buffer_load8(r, ok ? (in2 ? o2c[q] : o1c[q]) + srow : kBufferOOB);
        }
    };
    // biases of accumulator elements 4i..4i+3 (channels co0 + 8i + 4kh + 0..3)
    auto bias4 = [&](int i) __attribute__((always_inline)) { return *reinterpret_cast<const float4*>(s_bias + 8 * i + 4 * kh); };
    const bool relu = act == 1;
    const float neg_slope = act == 2 ? slope : 1.f;
    const int px = x0 + m;
    // OUT 0 / 1: the image's output (and pool) tensors as buffer resources
    const unsigned y_img = unsigned(H * W * Cout), p_img = unsigned((H / 2) * (W / 2) * Cout);
    const __amdgpu_buffer_rsrc_t ry = buffer_rsrc(OUT == 2 ? x1 : y + size_t(n) * y_img, OUT == 2 ? 0u : y_img * 2u);
    const __amdgpu_buffer_rsrc_t rp = buffer_rsrc(OUT == 1 ? ypool + size_t(n) * p_img : y, OUT == 1 ? p_img * 2u : 0u);
    // with OUT 1 (even rows per strip, even H and W) the 2x2 average pool of
    // the BF16 output as well, summed in the pooling kernel's order: the even
    // row's pixel pair, then the odd row's
    bf16x8 prev[2];
    // OUT 0 / 1: every row goes through here, `valid` false for the rows
    // outside the strip or the image
    auto store_tile_row = [&](int yy, const f32x16& acc, bool valid) __attribute__((always_inline)) {
        // the wave's reads of s_x for this row are issued before these
        // writes, its next chunk's writes after these reads
        __asm__ volatile("" ::: "memory");
        bf16x4 b[4];
        act16_bf16(acc, s_bias, kh, relu, neg_slope, b);
#pragma unroll
        for (int i = 0; i < 4; i++) *reinterpret_cast<bf16x4*>(s_x + m * OXP + 8 * i + 4 * kh) = b[i];
        __asm__ volatile("" ::: "memory");
        const bool odd = (yy & 1) != 0;
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int p = stage_col<4>(r, l), cg = l & 3;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(s_x + p * OXP + cg * 8);
            const bool ok = valid && x0 + p < W;
            buffer_store8(v, ry, ok ? unsigned(((yy * W + x0 + p) * Cout + co0 + cg * 8) * 2) : kBufferOOB);
            if constexpr (OUT == 1) {
                // the pool is computed on odd rows only (a wave-uniform
                // branch holding no global memory operation); its store is
                // issued on every row
                bf16x8 pool = v;
                if (odd) {
                    // pixel p + 1 of both rows sits sixteen lanes up
                    const u32x4 pa = __builtin_bit_cast(u32x4, prev[r]), pv = __builtin_bit_cast(u32x4, v);
                    u32x4 na, nv;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        na[j] = unsigned(__shfl_down(int(pa[j]), 16));
                        nv[j] = unsigned(__shfl_down(int(pv[j]), 16));
                    }
                    const bf16x8 bq = __builtin_bit_cast(bf16x8, na), d = __builtin_bit_cast(bf16x8, nv);
#pragma unroll
                    for (int j = 0; j < 8; j++)
                        pool[j] = bf16(0.25f * (float(prev[r][j]) + float(bq[j]) + float(v[j]) + float(d[j])));
                } else {
                    prev[r] = v;
                }
                const bool okp = ok && odd && (p & 1) == 0;
                buffer_store8(pool, rp,
                              okp ? unsigned((((yy >> 1) * (W >> 1) + ((x0 + p) >> 1)) * Cout + co0 + cg * 8) * 2)
                                  : kBufferOOB);
            }
        }
        __asm__ volatile("" ::: "memory");
    };
    // OUT 2
    auto store_row = [&](int yy, const f32x16& acc) __attribute__((always_inline)) {
        if (px >= W) return;
        if (rgb) {
            // the network's output layer: rgb = clamp(base + first three channels)
            if (co0 == 0 && kh == 0) {
                const size_t q = ((size_t(n) * H + yy) * W + px) * 3;
                const float4 bb = bias4(0);
                const float bj[3] = {bb.x, bb.y, bb.z};
#pragma unroll
                for (int j = 0; j < 3; j++)
                    rgb[q + j] = fminf(fmaxf(rbase[q + j] + act_fwd(acc[j] + bj[j], act, slope), 0.f), 1.f);
            }
            return;
        }
        const size_t o = ((size_t(n) * H + yy) * W + px) * Cout;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = co0 + 8 * i + 4 * kh;
            if (co >= Cout) continue;
            const float4 bb = bias4(i);
            const float bj[4] = {bb.x, bb.y, bb.z, bb.w};
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = act_select(acc[i * 4 + j] + bj[j], relu, neg_slope);
            if ((Cout & 3) == 0) {
                if (y) {
                    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                    bf16x4 b;
#pragma unroll
                    for (int j = 0; j < 4; j++) b[j] = bf16(v[j]);
                    *reinterpret_cast<bf16x4*>(y + o + co) = b;
                }
                if (yf) *reinterpret_cast<float4*>(yf + o + co) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (co + j >= Cout) break;
                    if (y) y[o + co + j] = bf16(v[j]);
                    if (yf) yf[o + co + j] = v[j];
                }
            }
        }
    };

    // Steps (input row jj, chunk ch) run jj-major; step s's chunk is loaded
    // at step s - 3 into register set s % 3. Rows are unrolled three at a
    // time, so both the set of a step and the roles of the accumulators are
    // fixed at compile time.
    bf16x8 pa[NQ], pb[NQ], pc[NQ];
    static_for<0, 3>([&](auto sc) __attribute__((always_inline)) {
        constexpr int S = decltype(sc)::value;
        auto& buf = [&]() __attribute__((always_inline)) -> bf16x8 (&)[NQ] {
            if constexpr (S == 0) return pa;
            else if constexpr (S == 1) return pb;
            else return pc;
        }();
        load_chunk(S / NCH, std::integral_constant<int, S % NCH>{}, buf, S / NCH < rows + 2);
    });
    // the tile's weights [32 out][9 taps][WP], their loads issued after the
    // first three chunks'
    stage_fragments<64 * NW, 9 * 32 * GW>(
        t,
        [&](int e) __attribute__((always_inline)) {
            const int g = e % GW, tap = (e / GW) % 9, co = e / (9 * GW);
            return co0 + co < Cout && g * 8 < Cin ? load8(wgt + unsigned(((co0 + co) * 9 + tap) * Cin + g * 8))
                                                  : zero8();
        },
        [&](int e, bf16x8 v) __attribute__((always_inline)) {
            const int g = e % GW, tap = (e / GW) % 9, co = e / (9 * GW);
            *reinterpret_cast<bf16x8*>(s_w + (co * 9 + tap) * WP + g * 8) = v;
        });
    if (t < 32) s_bias[t] = (bias && co0 + t < Cout) ? bias[co0 + t] : 0.f;
    __syncthreads();  // weights and biases staged; from here on every wave runs on its own
    if (x0 >= W) return;  // a wave wholly right of the image (the last column of workgroups)

    f32x16 a0, a1, a2;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        a0[i] = 0.f;
        a1[i] = 0.f;
        a2[i] = 0.f;
    }
    // one input row: aA/aB/aC accumulate output rows jj, jj - 1, jj - 2 (taps
    // dy = 0, 1, 2); afterwards row jj - 2 is complete
    auto row = [&](auto rc, int jj, f32x16& aA, f32x16& aB, f32x16& aC) __attribute__((always_inline)) {
        constexpr int R = decltype(rc)::value;
        const int gy = y0 - 1 + jj;
        const bool in_img = jj < rows + 2 && gy >= 0 && gy < H;
        // the halo rows feed only some of the three output rows: the MFMAs
        // of an accumulator whose row lies outside the strip are skipped
        const bool uA = jj < rows, uB = jj >= 1 && jj <= rows, uC = jj >= 2;
        static_for<0, NCH>([&](auto chc) __attribute__((always_inline)) {
            constexpr int CH = decltype(chc)::value, S = R * NCH + CH;
            auto& buf = [&]() __attribute__((always_inline)) -> bf16x8 (&)[NQ] {
                if constexpr (S % 3 == 0) return pa;
                else if constexpr (S % 3 == 1) return pb;
                else return pc;
            }();
            // the wave's previous reads of s_x are issued before these writes
            // (LDS executes one wave's instructions in order)
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int col = stage_col<G>(q, l);
                if (col < SPX) *reinterpret_cast<bf16x8*>(s_x + col * XP + g * 8) = buf[q];
            }
            __asm__ volatile("" ::: "memory");
            // step s + 3
            constexpr int NJ = (CH + 3) / NCH, NC = (CH + 3) % NCH;
            if constexpr (OUT == 2) {
                if (jj + NJ < rows + 2) load_chunk(jj + NJ, std::integral_constant<int, NC>{}, buf, true);
            } else {
                load_chunk(jj + NJ, std::integral_constant<int, NC>{}, buf, jj + NJ < rows + 2);
            }
            if (!in_img) return;  // rows outside the image or the strip add nothing
            const bf16* sx = s_x + m * XP + kh * 8;
            const bf16* sw = s_w + m * 9 * WP + CH * KC + kh * 8;
            // operands of step i = (dx = i / KS, ks = i % KS): the pixels are
            // read one step ahead, each weight fragment right after the MFMA
            // that used the previous one, three MFMAs before its own
            auto rx = [&](int i) __attribute__((always_inline)) {
                return *reinterpret_cast<const bf16x8*>(sx + (i / KS) * XP + (i % KS) * 16);
            };
            auto rw = [&](int i, int dy) __attribute__((always_inline)) {
                return *reinterpret_cast<const bf16x8*>(sw + (3 * dy + i / KS) * WP + (i % KS) * 16);
            };
            bf16x8 xc = rx(0), w0 = rw(0, 0), w1 = rw(0, 1), w2 = rw(0, 2);
            // a wave feeding the matrix core issues ahead of the waves
            // staging or storing (measured 3-6 % off every strip layer)
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int i = 0; i < NS; i++) {
                const bool more = i + 1 < NS;
                const bf16x8 xn = more ? rx(i + 1) : xc;
                __builtin_amdgcn_sched_barrier(0);
                if (uA) aA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, xc, aA, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w0 = rw(i + 1, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (uB) aB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, xc, aB, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w1 = rw(i + 1, 1);
                __builtin_amdgcn_sched_barrier(0);
                if (uC) aC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2, xc, aC, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w2 = rw(i + 1, 2);
                xc = xn;
            }
            __builtin_amdgcn_s_setprio(0);
        });
        if constexpr (OUT == 2) {
            if (jj >= 2 && y0 + jj - 2 < H) store_row(y0 + jj - 2, aC);
        } else {
            store_tile_row(y0 + jj - 2, aC, jj >= 2 && jj < rows + 2 && y0 + jj - 2 < H);
        }
#pragma unroll
        for (int i = 0; i < 16; i++) aC[i] = 0.f;
    };
    for (int jb = 0; jb < rows + 2; jb += 3) {
        static_for<0, 3>([&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            const int jj = jb + R;
            // OUT 0 / 1 run whole groups of three rows, the ones past the
            // strip as dead rows, so every group issues the same memory ops
            if constexpr (OUT == 2) {
                if (jj >= rows + 2) return;
            }
            if constexpr (R == 0) row(rc, jj, a0, a2, a1);
            else if constexpr (R == 1) row(rc, jj, a1, a0, a2);
            else row(rc, jj, a2, a1, a0);
        });
    }
}

// The two full-resolution convolutions at either end of the network in one
// pass, so the 32-channel activation between them never leaves the CU:
// A (Cin <= KA input channels -> 32, act_a) and then
//   BKIND 0: B (32 -> 32, act_b) written as BF16, with its 2x2 average pool
//            (the encoder's first level: features -> c0 -> c1 -> pool);
//   BKIND 1: the output layer (32 -> at most 8 channels, taps stacked as in
//            conv3x3_rgb8_kernel), rgb = clamp(base + act_b(first three))
//            (the decoder's end: c13 -> c14 -> rgb).
// Every wave owns a 32-column strip of A's output (image columns xs - 1 ..
// xs + 30, xs a multiple of 30) and produces B's output for the 30 columns
// xs .. xs + 29; it walks down its rows as the strip kernel does: an input
// row (fetched three rows ahead by buffer loads) feeds A's three rotating
// accumulators, A's completed row goes through LDS (zero outside the image,
// B's padding) into B's accumulators, and B's completed row is stored. The
// MFMA sequences are the strip kernel's (and the stacked output kernel's),
// so the results equal the two separate layers bit for bit.
template <int KA, int BKIND, int NW>
__global__ __launch_bounds__(64 * NW) void conv3x3_pair_kernel(const bf16* __restrict__ x, int Cin, int H, int W,
                                                           const bf16* __restrict__ wa, const float* __restrict__ ba,
                                                           int act_a, const bf16* __restrict__ wb,
                                                           const float* __restrict__ bb, int cout_b, int act_b,
                                                           int rows, bf16* __restrict__ y, bf16* __restrict__ ypool,
                                                           const float* __restrict__ rbase, float* __restrict__ rgb) {
    constexpr int SPX = 34, XP = KA + PAD, G = KA / 8, NQ = (SPX * G + 63) / 64, KSA = KA / 16, NSA = 3 * KSA;
    constexpr int WPA = KA + PAD, MP = 32 + PAD, OXP = 32 + PAD, XS = strip_wave_lds(KA);
    constexpr int WB_ROWS = BKIND == 0 ? 9 : 3;  // B's weight rows per output row of the MFMA tile
    static_assert(64 % G == 0, "a lane's channel group is the same in all its loads");
    extern __shared__ __attribute__((aligned(16))) unsigned char conv_smem[];
    const int t = int(threadIdx.x), wv = t >> 6, l = t & 63, m = l & 31, kh = l >> 5;
    // [32 out][9 taps][WPA]; B: [32 out][9 taps][MP] or [32 rows 8 dy + co][3 dx][MP]
    bf16* s_wa = reinterpret_cast<bf16*>(conv_smem);
    bf16* s_wb = s_wa + 32 * 9 * WPA;
    float* s_ba = reinterpret_cast<float*>(s_wb + 32 * WB_ROWS * MP);
    float* s_bb = s_ba + 32;
    bf16* s_x = reinterpret_cast<bf16*>(s_bb + 32) + wv * (XS + SPX * MP);  // staged input row / output row
    bf16* s_m = s_x + XS;                                                  // A's row [34][MP]
    const int xs = (int(blockIdx.x) * NW + wv) * 30;
    const int y0 = int(blockIdx.y) * rows;
    const int n = int(blockIdx.z);
    const bf16* xn = x + size_t(n) * H * W * Cin;
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(xn, unsigned(H * W * Cin) * 2u);

    if (t < 32) {
        s_ba[t] = ba ? ba[t] : 0.f;
        s_bb[t] = (bb && t < cout_b) ? bb[t] : 0.f;
    }
    // A's row columns 32 and 33 are B's operands only for the two unused lanes
    if (l < 2 * (MP / 8)) *reinterpret_cast<bf16x8*>(s_m + (32 + l / (MP / 8)) * MP + (l % (MP / 8)) * 8) = zero8();

    // the lane's part of a staged input row: channel group g, column stage_col(q, l) in round q
    const int g = l % G;
    unsigned oc[NQ];
    bool colok[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int col = stage_col<G>(q, l), gx = xs - 2 + col;
        colok[q] = col < SPX && gx >= 0 && gx < W;
        oc[q] = unsigned(gx * Cin + g * 8) * 2u;
    }
    const bool ch_ok = g * 8 < Cin;
    // input row jj = image row y0 - 2 + jj; zeros where `live` is false
    auto load_row = [&](int jj, bf16x8* v, bool live = true) __attribute__((always_inline)) {
        const int gy = y0 - 2 + jj;
        const bool row_ok = live && gy >= 0 && gy < H;
        const unsigned srow = unsigned(gy * W * Cin) * 2u;
#pragma unroll
        for (int q = 0; q < NQ; q++) v[q] = buffer_load8(rx, row_ok && colok[q] && ch_ok ? oc[q] + srow : kBufferOOB);
    };
    const int NJ = rows + 4;
    bf16x8 pa[NQ], pb[NQ], pc[NQ];
    load_row(0, pa);
    load_row(1, pb);
    load_row(2, pc);
    // both layers' weights, their loads issued after the first three rows'
    stage_fragments<64 * NW, 32 * 9 * (WPA / 8)>(
        t,
        [&](int e) __attribute__((always_inline)) {
            const int g = e % (WPA / 8), tap = (e / (WPA / 8)) % 9, co = e / (9 * (WPA / 8));
            return g * 8 < Cin ? load8(wa + unsigned((co * 9 + tap) * Cin + g * 8)) : zero8();
        },
        [&](int e, bf16x8 v) __attribute__((always_inline)) {
            const int g = e % (WPA / 8), tap = (e / (WPA / 8)) % 9, co = e / (9 * (WPA / 8));
            *reinterpret_cast<bf16x8*>(s_wa + (co * 9 + tap) * WPA + g * 8) = v;
        });
    stage_fragments<64 * NW, 32 * WB_ROWS * (MP / 8)>(
        t,
        [&](int e) __attribute__((always_inline)) {
            const int g = e % (MP / 8), j = (e / (MP / 8)) % WB_ROWS, r = e / (WB_ROWS * (MP / 8));
            bf16x8 v = zero8();
            if (g < 4) {
                if (BKIND == 0) {
                    if (r < cout_b) v = load8(wb + unsigned((r * 9 + j) * 32 + g * 8));
                } else {
                    const int dy = r >> 3, co = r & 7;
                    if (dy < 3 && co < cout_b) v = load8(wb + unsigned((co * 9 + dy * 3 + j) * 32 + g * 8));
                }
            }
            return v;
        },
        [&](int e, bf16x8 v) __attribute__((always_inline)) {
            const int g = e % (MP / 8), j = (e / (MP / 8)) % WB_ROWS, r = e / (WB_ROWS * (MP / 8));
            *reinterpret_cast<bf16x8*>(s_wb + (r * WB_ROWS + j) * MP + g * 8) = v;
        });
    __syncthreads();  // weights and biases staged; from here on every wave runs on its own
    if (xs >= W) return;

    const bool relu_a = act_a == 1, relu_b = act_b == 1;
    const float neg_a = act_a == 2 ? 0.f : 1.f, neg_b = act_b == 2 ? 0.f : 1.f;
    const int mx = xs - 1 + m;  // image column of A's output in this lane
    f32x16 a0, a1, a2, b0, b1, b2;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        a0[i] = 0.f;
        a1[i] = 0.f;
        a2[i] = 0.f;
        b0[i] = 0.f;
        b1[i] = 0.f;
        b2[i] = 0.f;
    }
    // the output layer: group 0 of A's two previous rows, group 1 of the previous one
    float q0a[3] = {0.f, 0.f, 0.f}, q0b[3] = {0.f, 0.f, 0.f}, q1a[3] = {0.f, 0.f, 0.f};
    bf16x8 prev[2];
    // Every row issues the same global loads and stores, dead ones through
    // buffer offsets beyond their resource (reads give zeros, writes are
    // dropped; a missing pool or output buffer is a resource of size 0), so
    // the wait counts before each staged row only wait for that row's loads.
    const unsigned y_img = unsigned(H * W * 32), p_img = unsigned((H / 2) * (W / 2) * 32), f_img = unsigned(H * W * 3);
    const void* any = x;  // the base of a resource of size 0
    const __amdgpu_buffer_rsrc_t ry = buffer_rsrc(y ? static_cast<const void*>(y + size_t(n) * y_img) : any, y ? y_img * 2u : 0u);
    const __amdgpu_buffer_rsrc_t rp =
        buffer_rsrc(ypool ? static_cast<const void*>(ypool + size_t(n) * p_img) : any, ypool ? p_img * 2u : 0u);
    const __amdgpu_buffer_rsrc_t rbr =
        buffer_rsrc(rbase ? static_cast<const void*>(rbase + size_t(n) * f_img) : any, rbase ? f_img * 4u : 0u);
    const __amdgpu_buffer_rsrc_t rrgb =
        buffer_rsrc(rgb ? static_cast<const void*>(rgb + size_t(n) * f_img) : any, rgb ? f_img * 4u : 0u);

    // B's output row yy (lane-transposed through s_x, as the strip kernel
    // stores), dropped where `valid` is false
    auto store_b = [&](int yy, const f32x16& acc, bool valid) __attribute__((always_inline)) {
        __asm__ volatile("" ::: "memory");
        bf16x4 b[4];
        act16_bf16(acc, s_bb, kh, relu_b, neg_b, b);
#pragma unroll
        for (int i = 0; i < 4; i++) *reinterpret_cast<bf16x4*>(s_x + m * OXP + 8 * i + 4 * kh) = b[i];
        __asm__ volatile("" ::: "memory");
        const bool odd = (yy & 1) != 0;
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int p = stage_col<4>(r, l), cg = l & 3;
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(s_x + p * OXP + cg * 8);
            const bool keep = valid && p < 30 && xs + p < W;
            buffer_store8(v, ry, keep ? unsigned(((yy * W + xs + p) * 32 + cg * 8) * 2) : kBufferOOB);
            // the pool (computed on odd rows, a branch without global memory
            // operations; its store issued on every row)
            bf16x8 pool = v;
            if (odd) {
                const u32x4 pu = __builtin_bit_cast(u32x4, prev[r]), vu = __builtin_bit_cast(u32x4, v);
                u32x4 nu, nv;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    nu[j] = unsigned(__shfl_down(int(pu[j]), 16));
                    nv[j] = unsigned(__shfl_down(int(vu[j]), 16));
                }
                const bf16x8 bq = __builtin_bit_cast(bf16x8, nu), d = __builtin_bit_cast(bf16x8, nv);
#pragma unroll
                for (int j = 0; j < 8; j++)
                    pool[j] = bf16(0.25f * (float(prev[r][j]) + float(bq[j]) + float(v[j]) + float(d[j])));
            } else {
                prev[r] = v;
            }
            const bool keep_pool = keep && odd && (p & 1) == 0;
            buffer_store8(pool, rp,
                          keep_pool ? unsigned((((yy >> 1) * (W >> 1) + ((xs + p) >> 1)) * 32 + cg * 8) * 2)
                                    : kBufferOOB);
        }
        __asm__ volatile("" ::: "memory");
    };

    // the output layer's pixel of input row jj's finished output row (image
    // row y0 + jj - 4): the byte offset of its three floats in base and rgb,
    // kBufferOOB where the lane has none
    auto out_off = [&](int jj) __attribute__((always_inline)) {
        const int oy = y0 + jj - 4, px = xs + m;
        const bool ok = BKIND == 1 && jj < NJ && jj >= 4 && oy < H && kh == 0 && m < 30 && px < W;
        return ok ? unsigned((oy * W + px) * 3) * 4u : kBufferOOB;
    };
    // its base pixels, fetched three rows before they are used
    float rba[3] = {0.f, 0.f, 0.f}, rbb[3] = {0.f, 0.f, 0.f}, rbc[3] = {0.f, 0.f, 0.f};
    auto load_base = [&](int jj, float (&v)[3]) __attribute__((always_inline)) {
        const unsigned o = out_off(jj);
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rbr, o + 4u * c, 0, 0));
    };
    if constexpr (BKIND == 1) {
        load_base(0, rba);
        load_base(1, rbb);
        load_base(2, rbc);
    }
    // input row jj: A's accumulators hold A rows jj, jj - 1, jj - 2 (aA, aB,
    // aC); B's hold B rows jj - 2, jj - 3, jj - 4 (bA, bB, bC), in the
    // relative numbering where A row q is image row y0 - 1 + q and B row o
    // image row y0 + o
    auto row = [&](auto rc, int jj, f32x16& aA, f32x16& aB, f32x16& aC, f32x16& bA, f32x16& bB, f32x16& bC)
        __attribute__((always_inline)) {
        constexpr int R = decltype(rc)::value;
        auto& buf = [&]() __attribute__((always_inline)) -> bf16x8 (&)[NQ] {
            if constexpr (R == 0) return pa;
            else if constexpr (R == 1) return pb;
            else return pc;
        }();
        auto& rb = [&]() __attribute__((always_inline)) -> float (&)[3] {
            if constexpr (R == 0) return rba;
            else if constexpr (R == 1) return rbb;
            else return rbc;
        }();
        const bool live = jj < NJ;  // rows past the strip's run as dead ones
        const int oy = y0 + jj - 4;
        __asm__ volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int col = stage_col<G>(q, l);
            if (col < SPX) *reinterpret_cast<bf16x8*>(s_x + col * XP + g * 8) = buf[q];
        }
        __asm__ volatile("" ::: "memory");
        load_row(jj + 3, buf, jj + 3 < NJ);
        const int gy = y0 - 2 + jj;
        if (live && gy >= 0 && gy < H) {
            // A rows outside 0 .. rows + 1 feed no B row of the strip
            const bool uA = jj < rows + 2, uB = jj >= 1 && jj <= rows + 2, uC = jj >= 2;
            const bf16* sx = s_x + m * XP + kh * 8;
            const bf16* sw = s_wa + m * 9 * WPA + kh * 8;
            auto rxa = [&](int i) __attribute__((always_inline)) {
                return *reinterpret_cast<const bf16x8*>(sx + (i / KSA) * XP + (i % KSA) * 16);
            };
            auto rwa = [&](int i, int dy) __attribute__((always_inline)) {
                return *reinterpret_cast<const bf16x8*>(sw + (3 * dy + i / KSA) * WPA + (i % KSA) * 16);
            };
            bf16x8 xc = rxa(0), w0 = rwa(0, 0), w1 = rwa(0, 1), w2 = rwa(0, 2);
            __builtin_amdgcn_s_setprio(1);  // as in the strip kernel
#pragma unroll
            for (int i = 0; i < NSA; i++) {
                const bool more = i + 1 < NSA;
                const bf16x8 xn2 = more ? rxa(i + 1) : xc;
                __builtin_amdgcn_sched_barrier(0);
                if (uA) aA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, xc, aA, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w0 = rwa(i + 1, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (uB) aB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, xc, aB, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w1 = rwa(i + 1, 1);
                __builtin_amdgcn_sched_barrier(0);
                if (uC) aC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2, xc, aC, 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (more) w2 = rwa(i + 1, 2);
                xc = xn2;
            }
            __builtin_amdgcn_s_setprio(0);
        }
        float vout[3] = {0.f, 0.f, 0.f};
        if (live && jj >= 2) {
            // A row q = jj - 2 (image row y0 - 3 + jj) is complete: activation,
            // BF16, zero outside the image, into s_m for B
            const int qy = y0 - 3 + jj;
            const bool ok = qy >= 0 && qy < H && mx >= 0 && mx < W;
            __asm__ volatile("" ::: "memory");
            bf16x4 b[4];
            act16_bf16(aC, s_ba, kh, relu_a, neg_a, b);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const bf16x4 v = ok ? b[i] : bf16x4{bf16(0.f), bf16(0.f), bf16(0.f), bf16(0.f)};
                *reinterpret_cast<bf16x4*>(s_m + m * MP + 8 * i + 4 * kh) = v;
            }
            __asm__ volatile("" ::: "memory");
            const bool q_in = qy >= 0 && qy < H;
            const bf16* sm = s_m + m * MP + kh * 8;
            auto rxb = [&](int i) __attribute__((always_inline)) {
                return *reinterpret_cast<const bf16x8*>(sm + (i / 2) * MP + (i % 2) * 16);
            };
            if constexpr (BKIND == 0) {
                if (q_in) {
                    // B rows outside the strip are skipped
                    const bool uA = jj < rows + 2, uB = jj >= 3 && jj < rows + 3, uC = jj >= 4;
                    const bf16* sw = s_wb + m * 9 * MP + kh * 8;
                    auto rwb = [&](int i, int dy) __attribute__((always_inline)) {
                        return *reinterpret_cast<const bf16x8*>(sw + (3 * dy + i / 2) * MP + (i % 2) * 16);
                    };
                    bf16x8 xc = rxb(0), w0 = rwb(0, 0), w1 = rwb(0, 1), w2 = rwb(0, 2);
                    __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        const bool more = i + 1 < 6;
                        const bf16x8 xn2 = more ? rxb(i + 1) : xc;
                        __builtin_amdgcn_sched_barrier(0);
                        if (uA) bA = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w0, xc, bA, 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                        if (more) w0 = rwb(i + 1, 0);
                        __builtin_amdgcn_sched_barrier(0);
                        if (uB) bB = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, xc, bB, 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                        if (more) w1 = rwb(i + 1, 1);
                        __builtin_amdgcn_sched_barrier(0);
                        if (uC) bC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2, xc, bC, 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                        if (more) w2 = rwb(i + 1, 2);
                        xc = xn2;
                    }
                    __builtin_amdgcn_s_setprio(0);
                }
            } else {
                // the stacked output layer: one accumulator per dx
                f32x16 acc[3];
#pragma unroll
                for (int d = 0; d < 3; d++)
#pragma unroll
                    for (int i = 0; i < 16; i++) acc[d][i] = 0.f;
                if (q_in) {
                    const bf16* sw = s_wb + m * 3 * MP + kh * 8;
                    __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int dx = 0; dx < 3; dx++)
#pragma unroll
                        for (int ks = 0; ks < 2; ks++) {
                            const bf16x8 xb = rxb(dx * 2 + ks);
                            const bf16x8 wv8 = *reinterpret_cast<const bf16x8*>(sw + dx * MP + ks * 16);
                            acc[dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv8, xb, acc[dx], 0, 0, 0);
                        }
                    __builtin_amdgcn_s_setprio(0);
                }
                float g0[3], g1[3], g2[3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    g0[c] = (acc[0][c] + acc[1][c]) + acc[2][c];
                    g1[c] = (acc[0][4 + c] + acc[1][4 + c]) + acc[2][4 + c];
                    g2[c] = (acc[0][8 + c] + acc[1][8 + c]) + acc[2][8 + c];
                }
                // B's output row image qy - 1 = y0 + jj - 4 is complete
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float v = q0b[c] + q1a[c] + g2[c];
                    vout[c] = fminf(fmaxf(rb[c] + act_fwd(v + s_bb[c], act_b, 0.f), 0.f), 1.f);
                }
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    q0b[c] = q0a[c];
                    q0a[c] = g0[c];
                    q1a[c] = g1[c];
                }
            }
        }
        if constexpr (BKIND == 0) {
            // B row o = jj - 4 (image row oy = y0 + jj - 4) is complete
            store_b(oy, bC, live && jj >= 4 && oy < H);
#pragma unroll
            for (int i = 0; i < 16; i++) bC[i] = 0.f;
        } else {
            const unsigned qo = out_off(jj);
#pragma unroll
            for (int c = 0; c < 3; c++)
                __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, vout[c]), rrgb, qo + 4u * c, 0, 0);
            // this set's next row, three rows on
            load_base(jj + 3, rb);
        }
        // A row jj - 2 is done (or, before row 2, outside the strip): aC
        // starts over as row jj + 1
#pragma unroll
        for (int i = 0; i < 16; i++) aC[i] = 0.f;
    };
    for (int jb = 0; jb < NJ; jb += 3) {
        static_for<0, 3>([&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            const int jj = jb + R;
            if constexpr (R == 0) row(rc, jj, a0, a2, a1, b0, b2, b1);
            else if constexpr (R == 1) row(rc, jj, a1, a0, a2, b1, b0, b2);
            else row(rc, jj, a2, a1, a0, b2, b1, b0);
        });
    }
}

// The network's output layer (at most 8 channels, rgb = clamp(base + act(
// channels 0..2 + bias))): the three row taps' weights are stacked into one
// 32-row MFMA tile, row 8 dy + co (rows 24..31 zero), so one MFMA per
// (dx, k step) does the work of three. The accumulator of input row r holds
// in register group dy (elements 4 dy + j, j = co for the lanes with kh = 0)
// its contribution to output row r + 1 - dy, so output row y is group 0 of
// input row y - 1 plus group 1 of row y plus group 2 of row y + 1, all in
// the same lane. Direct reads only, 32-channel chunks, NW waves per group.
template <int NW>
__global__ __launch_bounds__(64 * NW) void conv3x3_rgb8_kernel(const bf16* __restrict__ x, int Cin, int H, int W,
                                                               const bf16* __restrict__ wgt,
                                                               const float* __restrict__ bias, int Cout, int act,
                                                               int rows, const float* __restrict__ rbase,
                                                               float* __restrict__ rgb) {
    constexpr int KC = 32, SPX = 34, XP = KC + PAD, G = KC / 8, NQ = (SPX * G + 63) / 64;
    extern __shared__ __attribute__((aligned(16))) unsigned char conv_smem[];
    const int nch = (Cin + KC - 1) / KC, CinP = nch * KC, WP = CinP + PAD, GW = CinP / 8;
    const int t = int(threadIdx.x), wv = t >> 6, l = t & 63, m = l & 31, kh = l >> 5;
    bf16* s_w = reinterpret_cast<bf16*>(conv_smem);  // [3 dx][32 rows 8 dy + co][WP]
    bf16* s_x = s_w + 3 * 32 * WP + wv * SPX * XP;   // this wave's staged chunk [SPX][XP]
    const int x0 = int(blockIdx.x) * 32 * NW + wv * 32;
    const int y0 = int(blockIdx.y) * rows;
    const int n = int(blockIdx.z);
    const bf16* xn = x + size_t(n) * H * W * Cin;

    for (int e = t; e < 3 * 32 * GW; e += 64 * NW) {
        const int g = e % GW, r = (e / GW) % 32, dx = e / (32 * GW);
        const int dy = r >> 3, co = r & 7;
        bf16x8 v = zero8();
        if (dy < 3 && co < Cout && g * 8 < Cin) v = load8(wgt + unsigned((co * 9 + dy * 3 + dx) * Cin + g * 8));
        *reinterpret_cast<bf16x8*>(s_w + (dx * 32 + r) * WP + g * 8) = v;
    }
    auto load_chunk = [&](int jj, int ch, bf16x8* v) {
        const int gy = y0 - 1 + jj;
        const bool row_ok = gy >= 0 && gy < H;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = l + q * 64;
            const int g = e % G, col = e / G;
            const int gx = x0 - 1 + col, c = ch * KC + g * 8;
            v[q] = load_px8<0>(xn, Cin, nullptr, 0, H, W, gy, gx, c,
                               row_ok && col < SPX && gx >= 0 && gx < W && c < Cin);
        }
    };
    const float b0 = bias ? bias[0] : 0.f, b1 = bias && Cout > 1 ? bias[1] : 0.f, b2 = bias && Cout > 2 ? bias[2] : 0.f;
    const int px = x0 + m;
    bf16x8 nxt[NQ];
    load_chunk(0, 0, nxt);
    __syncthreads();  // weights staged; from here on every wave runs on its own

    // group 0 of the two previous input rows, group 1 of the previous one
    float q0a[3] = {0.f, 0.f, 0.f}, q0b[3] = {0.f, 0.f, 0.f}, q1a[3] = {0.f, 0.f, 0.f};
    for (int jj = 0; jj < rows + 2; jj++) {
        const int gy = y0 - 1 + jj;
        const bool in_img = gy >= 0 && gy < H;
        // one accumulator per dx tap: three independent MFMA chains
        f32x16 acc[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[d][i] = 0.f;
        for (int ch = 0; ch < nch; ch++) {
            // the wave's previous reads of s_x are issued before these writes
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int e = l + q * 64;
                const int g = e % G, col = e / G;
                if (col < SPX) *reinterpret_cast<bf16x8*>(s_x + col * XP + g * 8) = nxt[q];
            }
            __asm__ volatile("" ::: "memory");
            {
                int nj = jj, nc = ch + 1;
                if (nc == nch) {
                    nc = 0;
                    nj++;
                }
                if (nj < rows + 2) load_chunk(nj, nc, nxt);
            }
            if (!in_img) continue;
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
#pragma unroll
                for (int ks = 0; ks < KC / 16; ks++) {
                    const bf16x8 xb = *reinterpret_cast<const bf16x8*>(s_x + (m + dx) * XP + ks * 16 + kh * 8);
                    const bf16x8 wa =
                        *reinterpret_cast<const bf16x8*>(s_w + (dx * 32 + m) * WP + ch * KC + ks * 16 + kh * 8);
                    acc[dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, xb, acc[dx], 0, 0, 0);
                }
            }
        }
        float g0[3], g1[3], g2[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            g0[c] = (acc[0][c] + acc[1][c]) + acc[2][c];
            g1[c] = (acc[0][4 + c] + acc[1][4 + c]) + acc[2][4 + c];
            g2[c] = (acc[0][8 + c] + acc[1][8 + c]) + acc[2][8 + c];
        }
        // output row y0 + jj - 2 is complete: group 0 of input row jj - 2,
        // group 1 of row jj - 1, group 2 of this row
        const int yy = y0 + jj - 2;
        if (jj >= 2 && yy < H && kh == 0 && px < W) {
            const size_t q = ((size_t(n) * H + yy) * W + px) * 3;
            const float v0 = q0b[0] + q1a[0] + g2[0], v1 = q0b[1] + q1a[1] + g2[1], v2 = q0b[2] + q1a[2] + g2[2];
            rgb[q + 0] = fminf(fmaxf(rbase[q + 0] + act_fwd(v0 + b0, act, 0.f), 0.f), 1.f);
            rgb[q + 1] = fminf(fmaxf(rbase[q + 1] + act_fwd(v1 + b1, act, 0.f), 0.f), 1.f);
            rgb[q + 2] = fminf(fmaxf(rbase[q + 2] + act_fwd(v2 + b2, act, 0.f), 0.f), 1.f);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            q0b[c] = q0a[c];
            q0a[c] = g0[c];
            q1a[c] = g1[c];
        }
    }
}

// Weight-stationary strip: like the strip kernel, but the loop over 32-channel
// input chunks is outside and the chunk's weights for all nine taps sit in
// VGPRs (A operands, 18 fragments), while R output rows accumulate in AGPRs.
// Per staged input row the wave reads only the six input fragments from LDS
// (3 dx x 2 k-steps), each feeding up to three MFMAs (dy = 0, 1, 2). Input
// rows are staged once per chunk: (R + 2) / R of the input per chunk. MODE is
// x1's read (0 direct, 1 2x upsample, 2 2x2 average pool); element offsets
// are 32-bit (the host checks the tensors fit).
template <int R, int MODE>
__global__ __launch_bounds__(256, 2) void conv3x3_wstat_kernel(const bf16* __restrict__ x1, int c1,
                                                                            const bf16* __restrict__ x2, int c2,
                                                                            int H, int W,
                                                                            const bf16* __restrict__ wgt,
                                                                            const float* __restrict__ bias, int Cout,
                                                                            int act, float slope,
                                                                            bf16* __restrict__ y,
                                                                            float* __restrict__ yf) {
    constexpr int KC = 32, SPX = 34, XP = KC + PAD, G = KC / 8, NQ = (SPX * G + 63) / 64;
    __shared__ __attribute__((aligned(16))) bf16 s_x_all[4][SPX][XP];
    const int Cin = c1 + c2;
    const int nch = (Cin + KC - 1) / KC;
    const int t = int(threadIdx.x), wv = t >> 6, l = t & 63, m = l & 31, kh = l >> 5;
    bf16* s_x = &s_x_all[wv][0][0];
    const int x0 = int(blockIdx.x) * 128 + wv * 32;
    const int y0 = int(blockIdx.y) * R;
    const int ntile = (Cout + 31) / 32;
    const int n = int(blockIdx.z) / ntile, co0 = (int(blockIdx.z) % ntile) * 32;
    const int H1 = MODE == 1 ? H / 2 : (MODE == 2 ? 2 * H : H), W1 = MODE == 1 ? W / 2 : (MODE == 2 ? 2 * W : W);
    const bf16* x1n = x1 + size_t(n) * H1 * W1 * c1;
    const bf16* x2n = x2 ? x2 + size_t(n) * H * W * c2 : nullptr;

    auto load_chunk = [&](int jj, int ch, bf16x8* v) {
        const int gy = y0 - 1 + jj;
        const bool row_ok = gy >= 0 && gy < H;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int e = l + q * 64;
            const int g = e % G, col = e / G;
            const int gx = x0 - 1 + col, c = ch * KC + g * 8;
            // (branches, not the clamped branch-free load: measured faster
            // here, where whole rows and chunks fall outside more often)
            bf16x8 val = zero8();
            if (row_ok && col < SPX && gx >= 0 && gx < W && c < Cin) {
                if (c >= c1) {
                    val = load8(x2n + unsigned((gy * W + gx) * c2 + (c - c1)));
                } else if (MODE == 1) {
                    val = load8(x1n + unsigned(((gy >> 1) * W1 + (gx >> 1)) * c1 + c));
                } else if (MODE == 2) {
                    const unsigned r0 = unsigned(((2 * gy) * W1 + 2 * gx) * c1 + c), r1 = r0 + unsigned(W1 * c1);
                    val = pool8(x1n + r0, x1n + r0 + c1, x1n + r1, x1n + r1 + c1);
                } else {
                    val = load8(x1n + unsigned((gy * W + gx) * c1 + c));
                }
            }
            v[q] = val;
        }
    };

    f32x16 acc[R];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[r][i] = 0.f;
    // input rows are loaded two steps ahead (R + 2 is even, so the two
    // register sets keep their parity across chunks)
    static_assert((R & 1) == 0, "R must be even");
    bf16x8 pf[2][NQ];
    load_chunk(0, 0, pf[0]);
    load_chunk(1, 0, pf[1]);
    const int co_l = co0 + m;
    // whole 32-channel tiles on both sides (every layer of the network): the
    // weight loads need no per-lane bounds
    const bool full = (Cin & 31) == 0 && (Cout & 31) == 0;
    for (int ch = 0; ch < nch; ch++) {
        bf16x8 wf[9][2];
        if (full) {
            const bf16* wrow = wgt + unsigned(co_l * 9 * Cin + ch * KC + kh * 8);
#pragma unroll
            for (int tap = 0; tap < 9; tap++)
#pragma unroll
                for (int ks = 0; ks < 2; ks++) wf[tap][ks] = load8(wrow + tap * Cin + ks * 16);
        } else {
#pragma unroll
            for (int tap = 0; tap < 9; tap++)
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const int k = ch * KC + ks * 16 + kh * 8;
                    wf[tap][ks] =
                        (co_l < Cout && k < Cin) ? load8(wgt + unsigned((co_l * 9 + tap) * Cin + k)) : zero8();
                }
        }
#pragma unroll
        for (int jj = 0; jj < R + 2; jj++) {
            // the wave's reads of the previous row are issued before these
            // writes (LDS executes one wave's instructions in order)
            __asm__ volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int e = l + q * 64;
                const int g = e % G, col = e / G;
                if (col < SPX) *reinterpret_cast<bf16x8*>(s_x + col * XP + g * 8) = pf[jj & 1][q];
            }
            __asm__ volatile("" ::: "memory");
            if (jj + 2 < R + 2) load_chunk(jj + 2, ch, pf[jj & 1]);
            else if (ch + 1 < nch) load_chunk(jj + 2 - (R + 2), ch + 1, pf[jj & 1]);
            const int gy = y0 - 1 + jj;
            if (gy < 0 || gy >= H) continue;  // zero rows add nothing
#pragma unroll
            for (int dx = 0; dx < 3; dx++)
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const bf16x8 xb = *reinterpret_cast<const bf16x8*>(s_x + (m + dx) * XP + ks * 16 + kh * 8);
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int oo = jj - dy;
                        if (oo >= 0 && oo < R)
                            acc[oo] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[dy * 3 + dx][ks], xb, acc[oo], 0, 0, 0);
                    }
                }
        }
    }
    const int px = x0 + m;
    if (px >= W) return;
    float bv[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int co = co0 + 8 * i + 4 * kh + j;
            bv[i * 4 + j] = (bias && co < Cout) ? bias[co] : 0.f;
        }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int yy = y0 + r;
        if (yy >= H) break;
        const size_t o = ((size_t(n) * H + yy) * W + px) * Cout;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = co0 + 8 * i + 4 * kh;
            if (co >= Cout) continue;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = act_fwd(acc[r][i * 4 + j] + bv[i * 4 + j], act, slope);
            if ((Cout & 3) == 0) {
                if (y) {
                    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                    bf16x4 b;
#pragma unroll
                    for (int j = 0; j < 4; j++) b[j] = bf16(v[j]);
                    *reinterpret_cast<bf16x4*>(y + o + co) = b;
                }
                if (yf) *reinterpret_cast<float4*>(yf + o + co) = make_float4(v[0], v[1], v[2], v[3]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (co + j >= Cout) break;
                    if (y) y[o + co + j] = bf16(v[j]);
                    if (yf) yf[o + co + j] = v[j];
                }
            }
        }
    }
}

template <int R>
static void launch_wstat(const void* x1, int c1, int mode, const void* x2, int c2, int N, int H, int W,
                         const void* wgt, const float* bias, int Cout, int act, float slope, void* y, float* yf,
                         hipStream_t s) {
    dim3 grid((W + 127) / 128, (H + R - 1) / R, N * ((Cout + 31) / 32));
#define RFK_WS(M)                                                                                                 \
    conv3x3_wstat_kernel<R, M><<<grid, 256, 0, s>>>((const bf16*)x1, c1, (const bf16*)x2, c2, H, W, (const bf16*)wgt, \
                                                    bias, Cout, act, slope, (bf16*)y, yf)
    if (mode == 1) RFK_WS(1);
    else if (mode == 2) RFK_WS(2);
    else RFK_WS(0);
#undef RFK_WS
}

// the layer's weights for all nine taps, as the strip kernel keeps them in LDS
static size_t strip_weight_bytes(int Cin, int KC) {
    const int CinP = (Cin + KC - 1) / KC * KC;
    return size_t(9) * 32 * (CinP + PAD) * sizeof(bf16);
}

static size_t strip_lds_bytes(int Cin, int KC, int nw) {
    const int CinP = (Cin + KC - 1) / KC * KC;
    return (size_t(9) * 32 * (CinP + PAD) + 64 + size_t(nw) * strip_wave_lds(KC)) * sizeof(bf16);
}

// waves per strip workgroup (one workgroup per CU either way, its LDS being
// the layer's weights): direct reads 8, the upsampling read 12 where the
// weights leave room (96->32 at 1080p 153 -> 143 us), 10 where the row's
// 32-pixel strips come in tens but not in twelves (960 wide: three full
// workgroups instead of two and a half-empty one, 114 -> 110 us for
// 160->64 at 540x960), the pooling read 4
static int strip_waves(int mode, int Cin, int KC, int W = 0) {
    if (mode == 1) {
        const int strips = (W + 31) / 32;
        if (W > 0 && strips % 10 == 0 && strips % 12 != 0 && strip_lds_bytes(Cin, KC, 10) <= 160 * 1024) return 10;
        return strip_lds_bytes(Cin, KC, 12) <= 160 * 1024 ? 12 : 8;
    }
    return mode == 2 ? 4 : 8;
}

// Rows per strip: a workgroup's time grows with rows + 2 (the halo rows), and
// a launch takes ceil(workgroups / resident workgroups) rounds, so pick the
// row count that minimises rounds x (rows + 2); even counts only when the
// kernel also pools.
static int strip_rows(const void* fn, size_t lds, int nw, int zdim, int H, int W, bool even, int px = 32) {
    thread_local std::unordered_map<size_t, int> resident_cache;
    int resident = 0;
    const size_t key = reinterpret_cast<size_t>(fn) * 1000003u + lds;
    auto it = resident_cache.find(key);
    if (it != resident_cache.end()) {
        resident = it->second;
    } else {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * nw, lds) != hipSuccess)
            return 16;
        resident = resident_cache[key] = std::max(1, cus * per_cu);
    }
    const long cols = (W + px * nw - 1) / (px * nw);
    int best = 16;
    long best_cost = -1;
    for (int r = 2; r <= 64; r += even ? 2 : 1) {
        const long blocks = cols * ((H + r - 1) / r) * zdim;
        const long cost = (blocks + resident - 1) / resident * (r + 2);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = r;
        }
    }
    return best;
}

// one strip convolution's operands (see conv3x3_strip_kernel)
struct strip_args {
    const void* x1;
    int c1;
    const void* x2;
    int c2, N, H, W;
    const void* wgt;
    const float* bias;
    int Cout, act;
    float slope;
    void* y;
    void* ypool;
    float* yf;
    const float* rbase;
    float* rgb;
};

template <int KC, int NCH, int M, int NW, int OUT>
static hipError_t launch_strip_nch(const strip_args& a, size_t lds, hipStream_t s) {
    const void* fn = reinterpret_cast<const void*>(&conv3x3_strip_kernel<KC, NCH, M, NW, OUT>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    const int zdim = a.N * ((a.Cout + 31) / 32);
    const int rows = strip_rows(fn, lds, NW, zdim, a.H, a.W, a.ypool != nullptr);
    dim3 grid((a.W + 32 * NW - 1) / (32 * NW), (a.H + rows - 1) / rows, zdim);
    conv3x3_strip_kernel<KC, NCH, M, NW, OUT><<<grid, 64 * NW, lds, s>>>(
        (const bf16*)a.x1, a.c1, (const bf16*)a.x2, a.c2, a.H, a.W, (const bf16*)a.wgt, a.bias, a.Cout, a.act,
        a.slope, rows, (bf16*)a.y, (bf16*)a.ypool, a.yf, a.rbase, a.rgb);
    return hipGetLastError();
}

// the kernel instance for the layer's chunk count (KC 16: one chunk; KC 32:
// up to seven, 224 input channels, the most whose weights fit in LDS)
template <int KC, int M, int NW, int OUT>
static hipError_t launch_strip_mode(const strip_args& a, size_t lds, hipStream_t s) {
    const int nch = (a.c1 + a.c2 + KC - 1) / KC;
    if constexpr (KC == 16) {
        if (nch == 1) return launch_strip_nch<KC, 1, M, NW, OUT>(a, lds, s);
    } else {
        switch (nch) {
            case 1: return launch_strip_nch<KC, 1, M, NW, OUT>(a, lds, s);
            case 2: return launch_strip_nch<KC, 2, M, NW, OUT>(a, lds, s);
            case 3: return launch_strip_nch<KC, 3, M, NW, OUT>(a, lds, s);
            case 4: return launch_strip_nch<KC, 4, M, NW, OUT>(a, lds, s);
            case 5: return launch_strip_nch<KC, 5, M, NW, OUT>(a, lds, s);
            case 6: return launch_strip_nch<KC, 6, M, NW, OUT>(a, lds, s);
            case 7: return launch_strip_nch<KC, 7, M, NW, OUT>(a, lds, s);
            default: break;
        }
    }
    return hipErrorInvalidValue;
}

// the strip kernel's store (OUT, see conv3x3_strip_kernel) for a layer's outputs
static int strip_out(const strip_args& a) {
    if (a.y && !a.yf && !a.rgb && (a.Cout % 32) == 0) return a.ypool ? 1 : 0;
    return 2;
}

// the instances built: direct reads with every store; the upsampling and
// pooling reads with whole BF16 tiles and no pool (the other layers take the
// tiled kernel). The LDS size was computed for nw waves.
static bool strip_instance(int mode, int nw, int out) {
    if (mode == 0) return nw == 8;
    if (mode == 1) return (nw == 8 || nw == 10 || nw == 12) && out == 0;
    return mode == 2 && nw == 4 && out == 0;
}

template <int KC>
static hipError_t launch_strip(const strip_args& a, int mode, int nw, size_t lds, hipStream_t s) {
    const int out = strip_out(a);
    if (!strip_instance(mode, nw, out)) return hipErrorInvalidValue;
    if (mode == 1 && nw == 10) return launch_strip_mode<KC, 1, 10, 0>(a, lds, s);
    if (mode == 1) return nw == 12 ? launch_strip_mode<KC, 1, 12, 0>(a, lds, s) : launch_strip_mode<KC, 1, 8, 0>(a, lds, s);
    if (mode == 2) return launch_strip_mode<KC, 2, 4, 0>(a, lds, s);
    if (out == 0) return launch_strip_mode<KC, 0, 8, 0>(a, lds, s);
    if (out == 1) return launch_strip_mode<KC, 0, 8, 1>(a, lds, s);
    return launch_strip_mode<KC, 0, 8, 2>(a, lds, s);
}

template <int R, int C, int KCH>
static void launch_conv_fwd(const void* x1, int c1, int mode1, const void* x2, int c2, int N, int H, int W,
                            const void* wgt, const float* bias, int Cout, int act, float slope, void* y, float* yf,
                            hipStream_t s) {
    dim3 grid((W + 128 * R - 1) / (128 * R), H, N * ((Cout + 32 * C - 1) / (32 * C)));
    conv3x3_fwd_kernel<R, C, KCH><<<grid, 256, 0, s>>>((const bf16*)x1, c1, mode1, (const bf16*)x2, c2, N, H, W,
                                                       (const bf16*)wgt, bias, Cout, act, slope, (bf16*)y, yf);
}

hipError_t conv3x3_forward(const void* x1, int c1, const void* x2, int c2, int N, int H, int W, const void* wgt,
                           const float* bias, int Cout, int act, float slope, void* y, float* yf, hipStream_t s,
                           int x1_mode, void* ypool) {
    if ((c1 % 8) || (c2 % 8) || c1 <= 0 || c2 < 0 || Cout <= 0) return hipErrorInvalidValue;
    if (x1_mode < 0 || x1_mode > 2 || (x1_mode == 1 && ((H | W) & 1))) return hipErrorInvalidValue;
    if (ypool && (!y || (Cout % 8) || ((H | W) & 1))) return hipErrorInvalidValue;
    const int Cin = c1 + c2;
    // 32-bit offsets inside one image (the strip kernels' indexing): the
    // strip kernel's buffer reads take byte offsets below 2^31, the other
    // kernels element offsets
    const int H1 = x1_mode == 2 ? 2 * H : H, W1 = x1_mode == 2 ? 2 * W : W;
    const bool offsets32 = size_t(H1) * W1 * c1 < (size_t(1) << 30) && size_t(H) * W * std::max(c2, Cout) < (size_t(1) << 30) &&
                           size_t(Cout) * 9 * Cin < (size_t(1) << 31);
    // the input-stationary strip kernel wherever the layer's weights fit in
    // 140 KB of LDS next to the staging rows (inputs up to 224 channels):
    // measured fastest there (96->32 at 1080p 143 us against 192 us
    // weight-stationary, 128->128 at 135x240 against the tiled kernel)
    const int KC = Cin <= 16 ? 16 : 32;
    // the pool comes out of the store epilogue of whole 32-channel BF16
    // tiles read directly
    const bool pool_fused = ypool && !yf && (Cout % 32) == 0 && x1_mode == 0;
    const strip_args a{x1, c1, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, pool_fused ? ypool : nullptr, yf,
                       nullptr, nullptr};
    const int out = strip_out(a);
    const int nw = strip_waves(x1_mode, Cin, KC, W);
    const size_t lds = strip_lds_bytes(Cin, KC, nw);
    // (the strip kernel reads every KC-channel chunk from one of x1, x2)
    const bool chunks_split = c2 == 0 || c1 % KC == 0;
    if (strip_weight_bytes(Cin, KC) <= 140 * 1024 && lds <= 160 * 1024 && offsets32 && chunks_split) {
        if (strip_instance(x1_mode, nw, out)) {
            hipError_t e =
                KC == 16 ? launch_strip<16>(a, x1_mode, nw, lds, s) : launch_strip<32>(a, x1_mode, nw, lds, s);
            if (e != hipSuccess || !ypool || pool_fused) return e;
            return avgpool2_forward(y, N, H, W, Cout, ypool, s);
        }
    }
    // wider inputs: the weight-stationary strip; the tiled kernel for the rest
    if (Cin >= 96 && Cout <= 96 && offsets32) {
        launch_wstat<4>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
    } else if (Cout <= 32) {
        if (Cin <= 16)
            launch_conv_fwd<1, 1, 16>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
        else
            launch_conv_fwd<1, 1, 32>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
    } else if (Cout <= 64) {
        launch_conv_fwd<1, 2, 32>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
    } else if (Cout <= 96) {
        launch_conv_fwd<1, 3, 16>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
    } else {
        launch_conv_fwd<1, 4, 16>(x1, c1, x1_mode, x2, c2, N, H, W, wgt, bias, Cout, act, slope, y, yf, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !ypool) return e;
    return avgpool2_forward(y, N, H, W, Cout, ypool, s);
}

hipError_t conv3x3_forward_rgb(const void* x, int cin, int N, int H, int W, const void* wgt, const float* bias,
                               int Cout, int act, const float* base, float* out_rgb, hipStream_t s) {
    if ((cin % 8) || cin <= 0 || Cout < 3 || Cout > 32 || !base || !out_rgb) return hipErrorInvalidValue;
    if (Cout <= 8 && cin <= 96 && size_t(H) * W * cin < (size_t(1) << 31)) {
        // the stacked-tap output kernel (32->8 + rgb at 1080p inside a denoised
        // frame: 60 -> 53 us against the strip kernel)
        constexpr int NW = 16;
        const int CinP = (cin + 31) / 32 * 32;
        const size_t lds = (size_t(3) * 32 * (CinP + PAD) + size_t(NW) * 34 * (32 + PAD)) * sizeof(bf16);
        const void* fn = reinterpret_cast<const void*>(&conv3x3_rgb8_kernel<NW>);
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (e != hipSuccess) return e;
        const int rows = strip_rows(fn, lds, NW, N, H, W, false);
        dim3 grid((W + 32 * NW - 1) / (32 * NW), (H + rows - 1) / rows, N);
        conv3x3_rgb8_kernel<NW><<<grid, 64 * NW, lds, s>>>((const bf16*)x, cin, H, W, (const bf16*)wgt, bias, Cout, act,
                                                          rows, base, out_rgb);
        return hipGetLastError();
    }
    const int KC = cin <= 16 ? 16 : 32;
    const int nw = strip_waves(0, cin, KC);
    const size_t lds = strip_lds_bytes(cin, KC, nw);
    if (strip_weight_bytes(cin, KC) > 140 * 1024 || lds > 160 * 1024 ||
        size_t(H) * W * std::max(cin, 3) >= (size_t(1) << 30))
        return hipErrorInvalidValue;
    const strip_args a{x, cin, nullptr, 0, N, H, W, wgt, bias, Cout, act, 0.f, nullptr, nullptr, nullptr, base, out_rgb};
    return KC == 16 ? launch_strip<16>(a, 0, nw, lds, s) : launch_strip<32>(a, 0, nw, lds, s);
}

template <int KA, int BKIND>
static hipError_t launch_pair(const void* x, int cin, int N, int H, int W, const void* wa, const float* ba, int act_a,
                              const void* wb, const float* bb, int cout_b, int act_b, void* y, void* ypool,
                              const float* base, float* out_rgb, hipStream_t s) {
    constexpr int NW = 8, MP = 32 + PAD;
    const size_t lds = (size_t(32) * 9 * (KA + PAD) + size_t(32) * (BKIND == 0 ? 9 : 3) * MP + 128 +
                        size_t(NW) * (strip_wave_lds(KA) + 34 * MP)) *
                       sizeof(bf16);
    const void* fn = reinterpret_cast<const void*>(&conv3x3_pair_kernel<KA, BKIND, NW>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    const int rows = strip_rows(fn, lds, NW, N, H, W, ypool != nullptr, 30);
    dim3 grid((W + 30 * NW - 1) / (30 * NW), (H + rows - 1) / rows, N);
    conv3x3_pair_kernel<KA, BKIND, NW><<<grid, 64 * NW, lds, s>>>((const bf16*)x, cin, H, W, (const bf16*)wa, ba,
                                                                   act_a, (const bf16*)wb, bb, cout_b, act_b, rows,
                                                                   (bf16*)y, (bf16*)ypool, base, out_rgb);
    return hipGetLastError();
}

hipError_t conv3x3_pair_forward(const void* x, int cin, int N, int H, int W, const void* wa, const float* ba,
                                int act_a, const void* wb, const float* bb, int cout_b, int act_b, void* y,
                                void* ypool, const float* base, float* out_rgb, hipStream_t s) {
    if (!x || !wa || !wb || cin <= 0 || cin > 32 || (cin % 8) || N <= 0 || H <= 0 || W <= 0) return hipErrorInvalidValue;
    if (size_t(H) * W * 32 * 2 >= (size_t(1) << 30)) return hipErrorInvalidValue;
    const bool rgb_out = out_rgb != nullptr;
    if (rgb_out == (y != nullptr)) return hipErrorInvalidValue;
    if (rgb_out && (!base || cout_b < 3 || cout_b > 8 || ypool)) return hipErrorInvalidValue;
    if (!rgb_out && (cout_b != 32 || (ypool && ((H | W) & 1)))) return hipErrorInvalidValue;
    if (cin <= 16)
        return rgb_out ? launch_pair<16, 1>(x, cin, N, H, W, wa, ba, act_a, wb, bb, cout_b, act_b, y, ypool, base,
                                            out_rgb, s)
                       : launch_pair<16, 0>(x, cin, N, H, W, wa, ba, act_a, wb, bb, cout_b, act_b, y, ypool, base,
                                            out_rgb, s);
    return rgb_out ? launch_pair<32, 1>(x, cin, N, H, W, wa, ba, act_a, wb, bb, cout_b, act_b, y, ypool, base, out_rgb,
                                        s)
                   : launch_pair<32, 0>(x, cin, N, H, W, wa, ba, act_a, wb, bb, cout_b, act_b, y, ypool, base, out_rgb,
                                        s);
}

size_t conv3x3_wgrad_workspace(int c1, int c2, int Cout, int N, int H, int W, int* groups, int* tiles_per_group) {
    const long tiles = long(N) * H * ((W + WG_PX - 1) / WG_PX);
    int g = int(std::min<long>(tiles, 256));
    int tpg = int((tiles + g - 1) / g);
    g = int((tiles + tpg - 1) / tpg);
    if (groups) *groups = g;
    if (tiles_per_group) *tiles_per_group = tpg;
    return size_t(g) * Cout * 9 * (c1 + c2) * sizeof(float);
}

hipError_t conv3x3_wgrad(const void* x1, int c1, const void* x2, int c2, const void* dy, int Cout, int N, int H, int W,
                         float* dw, float beta, float* workspace, hipStream_t s) {
    if ((c1 % 8) || (c2 % 8) || (Cout % 8)) return hipErrorInvalidValue;
    int groups = 0, tpg = 0;
    conv3x3_wgrad_workspace(c1, c2, Cout, N, H, W, &groups, &tpg);
    const int Cin = c1 + c2;
    dim3 grid(groups, (Cout + 31) / 32, (Cin + 31) / 32);
    conv3x3_wgrad_kernel<<<grid, 256, 0, s>>>((const bf16*)x1, c1, (const bf16*)x2, c2, (const bf16*)dy, Cout, N, H, W,
                                              tpg, workspace);
    const size_t count = size_t(Cout) * 9 * Cin;
    reduce_partials_kernel<<<nb(count), 256, 0, s>>>(workspace, groups, count, dw, beta);
    return hipGetLastError();
}

hipError_t bias_grad(const void* dy, size_t npix, int C, float* db, float beta, hipStream_t s) {
    bias_grad_kernel<<<C, 256, 0, s>>>((const bf16*)dy, npix, C, db, beta);
    return hipGetLastError();
}

hipError_t flip_weights(const void* w, int Cout, int Cin, void* wt, hipStream_t s) {
    size_t n = size_t(Cout) * 9 * Cin;
    flip_weights_kernel<<<nb(n), 256, 0, s>>>((const bf16*)w, Cout, Cin, (bf16*)wt);
    return hipGetLastError();
}

hipError_t act_backward(const void* y, const void* dy, size_t n, int act, float slope, void* dx, hipStream_t s) {
    act_bwd_kernel<<<nb(n), 256, 0, s>>>((const bf16*)y, (const bf16*)dy, n, act, slope, (bf16*)dx);
    return hipGetLastError();
}

hipError_t avgpool2_forward(const void* x, int N, int H, int W, int C, void* y, hipStream_t s) {
    size_t n = size_t(N) * (H / 2) * (W / 2) * C;
    if (C % 8) return hipErrorInvalidValue;
    avgpool2_fwd_kernel<<<nb(n / 8), 256, 0, s>>>((const bf16*)x, N, H, W, C, (bf16*)y);
    return hipGetLastError();
}

hipError_t avgpool2_backward(const void* dy, int N, int H, int W, int C, void* dx, hipStream_t s) {
    size_t n = size_t(N) * H * W * C;
    avgpool2_bwd_kernel<<<nb(n), 256, 0, s>>>((const bf16*)dy, N, H, W, C, (bf16*)dx);
    return hipGetLastError();
}

hipError_t upsample2_forward(const void* x, int N, int H, int W, int C, void* y, hipStream_t s) {
    size_t n = size_t(N) * H * 2 * W * 2 * C;
    upsample2_fwd_kernel<<<nb(n), 256, 0, s>>>((const bf16*)x, N, H, W, C, (bf16*)y);
    return hipGetLastError();
}

hipError_t upsample2_backward(const void* dy, int N, int H, int W, int C, void* dx, hipStream_t s) {
    size_t n = size_t(N) * H * W * C;
    upsample2_bwd_kernel<<<nb(n), 256, 0, s>>>((const bf16*)dy, N, H, W, C, (bf16*)dx);
    return hipGetLastError();
}

hipError_t split_add(const void* g, size_t npix, int c1, int c2, void* dx1, void* dx2, hipStream_t s) {
    size_t n = npix * size_t(c1 + c2);
    split_add_kernel<<<nb(n), 256, 0, s>>>((const bf16*)g, npix, c1, c2, (bf16*)dx1, (bf16*)dx2);
    return hipGetLastError();
}

hipError_t add_bf16(const void* a, void* b, size_t n, hipStream_t s) {
    add_bf16_kernel<<<nb(n), 256, 0, s>>>((const bf16*)a, (bf16*)b, n);
    return hipGetLastError();
}

hipError_t residual_out(const float* base, int Cb, const float* r, int Cr, size_t npix, float* out, hipStream_t s) {
    residual_out_kernel<<<nb(npix * 3), 256, 0, s>>>(base, Cb, r, Cr, npix, out);
    return hipGetLastError();
}

hipError_t mse_grad(const float* out, const float* target, size_t npix, int Cr, float scale, void* dr, float* loss_acc,
                    hipStream_t s) {
    mse_grad_kernel<<<nb(npix * Cr), 256, 0, s>>>(out, target, npix, Cr, scale, (bf16*)dr, loss_acc);
    return hipGetLastError();
}

hipError_t adam(float* w, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps,
                int step, float wd, void* wb, hipStream_t s) {
    float bc1 = 1.f - powf(b1, float(step)), bc2 = 1.f - powf(b2, float(step));
    adam_kernel<<<nb(n), 256, 0, s>>>(w, m, v, g, n, lr, b1, b2, eps, bc1, bc2, wd, (bf16*)wb);
    return hipGetLastError();
}

hipError_t f32_to_bf16(const float* a, void* b, size_t n, hipStream_t s) {
    f32_to_bf16_kernel<<<nb(n), 256, 0, s>>>(a, (bf16*)b, n);
    return hipGetLastError();
}

hipError_t bf16_to_f32(const void* a, float* b, size_t n, hipStream_t s) {
    bf16_to_f32_kernel<<<nb(n), 256, 0, s>>>((const bf16*)a, b, n);
    return hipGetLastError();
}

hipError_t features(const uint32_t* hist, bool hist64, const float* tm_rgba, size_t npix, int Cpad, void* feat,
                    float* base,
                    hipStream_t s) {
    features_kernel<<<nb(npix), 256, 0, s>>>(hist, hist64, (const float4*)tm_rgba, npix, Cpad, (bf16*)feat, base);
    return hipGetLastError();
}

hipError_t gather_crops(const void* src, int H, int W, int pixel_bytes, const int* offs, int n, int ch, int cw,
                        void* dst, hipStream_t s) {
    dim3 grid(std::max(1, std::min(64, (cw * pixel_bytes + 255) / 256)), ch, n);
    gather_crops_kernel<<<grid, 256, 0, s>>>((const unsigned char*)src, H, W, pixel_bytes, offs, ch, cw,
                                             (unsigned char*)dst);
    return hipGetLastError();
}

hipError_t image_features(const void* img, int fmt, int W, int H, int Wp, int Hp, int Cpad, void* feat, float* base,
                          hipStream_t s) {
    image_features_kernel<<<nb(size_t(Wp) * Hp), 256, 0, s>>>(img, fmt, W, H, Wp, Hp, Cpad, (bf16*)feat, base);
    return hipGetLastError();
}

hipError_t image_output(const float* rgb, int Wp, const void* img, int in_fmt, void* out, int out_fmt, int W, int H,
                        hipStream_t s) {
    image_output_kernel<<<nb(size_t(W) * H), 256, 0, s>>>(rgb, Wp, img, in_fmt, out, out_fmt, W, H);
    return hipGetLastError();
}

hipError_t clamp_rgb(const float* in, size_t n, float* out, hipStream_t s) {
    clamp_rgb_kernel<<<nb(n), 256, 0, s>>>(in, n, out);
    return hipGetLastError();
}

}  // namespace rfk::nn
