// Host launchers of the denoiser kernels (unet.hip). Activations are NHWC
// BF16, weights [Cout][9][Cin] BF16, accumulation FP32.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace rfk::nn {

// y = act(conv3x3(concat(x1', x2)) + bias); c1, c2 multiples of 8 (c2 may be 0).
// act: 0 none, 1 ReLU, 2 leaky ReLU(slope). y (BF16) and/or yf (FP32) may be null.
// x1_mode: 0 x1' = x1 (H x W), 1 x1' = nearest 2x upsample of x1 (H/2 x W/2),
// 2 x1' = 2x2 average pool of x1 (2H x 2W). ypool (needs y, even H and W,
// Cout a multiple of 8): also the 2x2 average pool of y, N x H/2 x W/2 x Cout
// BF16, as avgpool2_forward computes it.
hipError_t conv3x3_forward(const void* x1, int c1, const void* x2, int c2, int N, int H, int W, const void* wgt,
                           const float* bias, int Cout, int act, float slope, void* y, float* yf, hipStream_t s,
                           int x1_mode = 0, void* ypool = nullptr);
// The inference network's output layer with the residual fused:
// out_rgb[p][c] = clamp(base[p][c] + conv(x)[p][c], 0, 1) for c < 3 (FP32,
// N*H*W*3). hipErrorInvalidValue when the strip kernel does not take the shape.
hipError_t conv3x3_forward_rgb(const void* x, int cin, int N, int H, int W, const void* wgt, const float* bias,
                               int Cout, int act, const float* base, float* out_rgb, hipStream_t s);
// Two full-resolution layers in one pass: A = act_a(conv3x3(x) + ba), x with
// cin <= 32 channels, A 32 channels (never stored), then either
//   y (BF16, 32 channels) = act_b(conv3x3(A) + bb), cout_b = 32, and with
//   ypool its 2x2 average pool (even H and W); or
//   out_rgb = clamp(base + act_b(conv3x3(A) + bb)[:3], 0, 1), 3 <= cout_b <= 8.
// Equal bit for bit to conv3x3_forward twice (and conv3x3_forward_rgb);
// hipErrorInvalidValue for other shapes.
hipError_t conv3x3_pair_forward(const void* x, int cin, int N, int H, int W, const void* wa, const float* ba,
                                int act_a, const void* wb, const float* bb, int cout_b, int act_b, void* y,
                                void* ypool, const float* base, float* out_rgb, hipStream_t s);
// bytes of FP32 workspace conv3x3_wgrad needs
size_t conv3x3_wgrad_workspace(int c1, int c2, int Cout, int N, int H, int W, int* groups, int* tiles_per_group);
// dw = beta * dw + dL/dW (FP32, [Cout][9][Cin])
hipError_t conv3x3_wgrad(const void* x1, int c1, const void* x2, int c2, const void* dy, int Cout, int N, int H, int W,
                         float* dw, float beta, float* workspace, hipStream_t s);
hipError_t bias_grad(const void* dy, size_t npix, int C, float* db, float beta, hipStream_t s);
// wt[ci][tap][co] = w[co][8 - tap][ci] (data gradient = forward conv of dY with wt)
hipError_t flip_weights(const void* w, int Cout, int Cin, void* wt, hipStream_t s);
hipError_t act_backward(const void* y, const void* dy, size_t n, int act, float slope, void* dx, hipStream_t s);
hipError_t avgpool2_forward(const void* x, int N, int H, int W, int C, void* y, hipStream_t s);
hipError_t avgpool2_backward(const void* dy, int N, int H, int W, int C, void* dx, hipStream_t s);
hipError_t upsample2_forward(const void* x, int N, int H, int W, int C, void* y, hipStream_t s);
hipError_t upsample2_backward(const void* dy, int N, int H, int W, int C, void* dx, hipStream_t s);
hipError_t split_add(const void* g, size_t npix, int c1, int c2, void* dx1, void* dx2, hipStream_t s);
hipError_t add_bf16(const void* a, void* b, size_t n, hipStream_t s);
hipError_t residual_out(const float* base, int Cb, const float* r, int Cr, size_t npix, float* out, hipStream_t s);
hipError_t mse_grad(const float* out, const float* target, size_t npix, int Cr, float scale, void* dr, float* loss_acc,
                    hipStream_t s);
hipError_t adam(float* w, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps,
                int step, float wd, void* wb, hipStream_t s);
hipError_t f32_to_bf16(const float* a, void* b, size_t n, hipStream_t s);
hipError_t bf16_to_f32(const void* a, float* b, size_t n, hipStream_t s);
hipError_t features(const uint32_t* hist, bool hist64, const float* tm_rgba, size_t npix, int Cpad, void* feat,
                    float* base,
                    hipStream_t s);
hipError_t clamp_rgb(const float* in, size_t n, float* out, hipStream_t s);
// features (Cpad BF16 channels) and FP32 RGB base, Wp x Hp (padded with empty
// pixels), from a tonemapped image of W x H pixels in format fmt (0 float4,
// 1 half4, 2 float3, 3 half3; the alpha of a 4-channel image gives the
// density)
hipError_t image_features(const void* img, int fmt, int W, int H, int Wp, int Hp, int Cpad, void* feat, float* base,
                          hipStream_t s);
// rgb (Wp-wide FP32 rows) into a W x H image of out_fmt, alpha from img
hipError_t image_output(const float* rgb, int Wp, const void* img, int in_fmt, void* out, int out_fmt, int W, int H,
                        hipStream_t s);
// crops (img, y, x) from a stack of H x W images with pixel_bytes per pixel
hipError_t gather_crops(const void* src, int H, int W, int pixel_bytes, const int* offs, int n, int ch, int cw,
                        void* dst, hipStream_t s);

}  // namespace rfk::nn
