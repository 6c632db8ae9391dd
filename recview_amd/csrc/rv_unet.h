// rv_unet.h -- kernels of the learned restoration stage (residual U-Net on
// reconstructed slices).  Activations are NHWC bf16 (stored as uint16_t), the
// 3x3 convolutions run on the gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16) with
// FP32 accumulation; the first (1 -> C) and last (C -> 1) convolutions and the
// pooling / upsampling / loss / optimiser steps are VALU kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rv {

// Y[p][co] = act(bias[co] + sum_{tap,ci} W[co][tap][ci] * X[p + d(tap)][ci]),
// d(tap) = (tap/3 - 1, tap%3 - 1), zero outside the image.  Pixels p run over
// B x H x W; tensors are addressed as base + pixel*stride + off + channel.
// wfrag holds W in MFMA fragment order (pack_conv_weights).  mask (optional)
// zeroes outputs whose mask value is not > 0 (ReLU derivative of the input of
// the previous layer, used when this launch computes a data gradient).
struct ConvArgs {
    const uint16_t* in; int in_stride, in_off, Cin;
    const uint16_t* wfrag;
    const float* bias;            // nullptr: no bias
    uint16_t* out; int out_stride, out_off, Cout;
    const uint16_t* mask; int mask_stride, mask_off;   // nullptr: no mask
    int relu;
    int B, H, W;
    // Optional fusions (nullptr: off).
    // up: input channels [0, Cup) are read from this H/2 x W/2 tensor (Cup
    //     channels, contiguous) at pixel (y/2, x/2) -- nearest 2x upsampling of
    //     the concat's first part; channels [Cup, Cin) come from `in` as usual.
    const uint16_t* up; int Cup;
    // pool_out: also writes the 2x2 average of the (bf16) outputs, H/2 x W/2 x
    //     Cout contiguous (H and W even).
    uint16_t* pool_out;
};
hipError_t launch_conv3x3(const ConvArgs& a, hipStream_t s);
// The edge convolution (1 -> 32, weights ew [32][3][3], biases eb, ReLU) of the
// FP32 image x (B x H x W) fused into the 32 -> 32 convolution a that reads its
// output: a.in is not read, a.Cin = a.Cout = 32, no mask and no upsampled input.
// Bit-identical to launch_edge_in followed by launch_conv3x3.
hipError_t launch_edge_conv3x3_c32(const ConvArgs& a, const float* x, const float* ew, const float* eb,
                                   hipStream_t s);

// Weight gradient of the same convolution: for every split s (grid z) writes
//   part[s][co][tap][ci] = sum over the split's pixels of dY[p][co] * X[p + d(tap)][ci]
//   partb[s][co]         = sum over the split's pixels of dY[p][co]
// The caller sums the splits (launch_sum_splits).  Returns the split count.
struct WgradArgs {
    const uint16_t* dy; int dy_stride, dy_off, Cout;
    const uint16_t* x; int x_stride, x_off, Cin;
    float* part; float* partb;
    int B, H, W;
    int nsplit;
};
int wgrad_splits(int Cin, int Cout, int B, int H, int W);
hipError_t launch_conv3x3_wgrad(const WgradArgs& a, hipStream_t s);
// out[i] = sum_s part[s*n + i], i < n.
hipError_t launch_sum_splits(const float* part, int nsplit, int n, float* out, hipStream_t s);

// Packs FP32 master weights W[Cout][9][Cin] into fragment order, for the
// forward convolution (transpose = 0) or for its data gradient (transpose = 1:
// W'[ci][tap][co] = W[co][8 - tap][ci]).
hipError_t launch_pack_conv_weights(const float* w, int Cin, int Cout, int transpose,
                                    uint16_t* wfrag, hipStream_t s);

// First layer: 1 -> C channels (fp32 input image), ReLU, bf16 out.
hipError_t launch_edge_in(const float* x, const float* w, const float* b, int C, uint16_t* out,
                          int out_stride, int out_off, int B, int H, int W, hipStream_t s);
// Its weight gradient: part[blk] = weights [c][tap] (9*C) then biases (C).
int edge_blocks(int B, int H, int W);
hipError_t launch_edge_in_wgrad(const uint16_t* dy, int C, const float* x, float* part, int nblk,
                                int B, int H, int W, hipStream_t s);
// Last layer: y[p] = xres[p] + b + sum W[tap][c] * X[p + d(tap)][c]  (C -> 1, fp32 out).
hipError_t launch_edge_out(const uint16_t* x, int C, const float* w, const float* b,
                           const float* xres, float* y, int B, int H, int W, hipStream_t s);
// dX[p][c] = (X[p][c] > 0) * sum_tap W[tap][c] * dy[p - d(tap)]
hipError_t launch_edge_out_bwd(const float* dy, const float* w, const uint16_t* x, int C,
                               uint16_t* dx, int B, int H, int W, hipStream_t s);
// part[blk][tap*C + c] (9*C entries) and part[blk][9*C] = sum dy.
hipError_t launch_edge_out_wgrad(const float* dy, const uint16_t* x, int C, float* part, int nblk,
                                 int B, int H, int W, hipStream_t s);

// 2x2 average pooling (H, W: input size), nearest 2x upsampling into a
// channel slice of a wider tensor (H, W: output size) and its gradient.
hipError_t launch_avgpool2(const uint16_t* in, int in_stride, int in_off, int C, uint16_t* out,
                           int B, int H, int W, hipStream_t s);
hipError_t launch_upsample2(const uint16_t* in, int C, uint16_t* out, int out_stride, int out_off,
                            int B, int H, int W, hipStream_t s);
hipError_t launch_upsample2_bwd(const uint16_t* g, int g_stride, int g_off, int C, uint16_t* out,
                                int B, int H, int W, hipStream_t s);
// out[p][c] = (skip[p][c] + 0.25 * dp[p/2][c]) * (e[p][c] > 0)   (H, W: full size)
hipError_t launch_pool_bwd_add(const uint16_t* skip, int skip_stride, int skip_off,
                               const uint16_t* dp, const uint16_t* e, int e_stride, int e_off,
                               int C, uint16_t* out, int B, int H, int W, hipStream_t s);

// y = x * scale (n floats)
hipError_t launch_scale(const float* x, float scale, float* y, long long n, hipStream_t s);
// dy = 2 (y - t) / n ; part[blk] = sum (y - t)^2 ; returns the block count.
int loss_blocks(long long n);
hipError_t launch_mse(const float* y, const float* t, float* dy, float* part, long long n,
                      hipStream_t s);
// Adam with bias correction on n parameters.
hipError_t launch_adam(float* p, const float* g, float* m, float* v, long long n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s);

}  // namespace rv
