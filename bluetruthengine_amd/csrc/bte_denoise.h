// Learned denoiser: a 4-level U-Net (32/64/128/128 channels, 3x3 convolutions,
// skip connections) over 9 guide channels {log(1+RGB), albedo, normal}.  It
// fills the reference's denoiser slot (src/host/denoiser.cu:72-109, the
// closed OptiX HDR denoiser) with hand-written kernels: every convolution is an
// implicit GEMM on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32
// accumulation; backward data reuses the forward kernel with transposed,
// flipped weights; backward weights is a split-K MFMA reduction.
//
// Tensors are NHWC bf16; channel counts are multiples of 32 (the 9 input
// channels are zero-padded to 32) so every 8-channel lane load is a 16-byte
// vector.  Spatial sizes are multiples of 8 (three 2x2 poolings).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bte {
namespace dn {

typedef __bf16 bf16;

constexpr int kInChannels = 9;     // log(1+RGB), albedo, normal
constexpr int kInPad = 32;         // stored input channels
constexpr int kOutChannels = 3;
// Input firefly suppression: a pixel whose luminance exceeds this multiple of
// the largest luminance among its in-image 8 neighbours is scaled down to it.
constexpr float kFireflyRatio = 6.0f;

// One 3x3 convolution (stride 1, zero padding 1):
//   dst[p][co] = epilogue( sum_{tap, ci} src(p + tap)[ci] * w[co][tap][ci] )
// src = channels [0, c0) of src0 (optionally at half resolution, read through
// a nearest 2x upsample) followed by channels [c0, c0 + c1) of src1.
struct ConvArgs {
    const bf16* src0; int c0; int up0;
    const bf16* src1; int c1;
    const bf16* w;                 // [coutPad][9][c0 + c1]
    const float* bias;             // [cout] or null
    int n, h, w_;                  // output size (= full-resolution input size)
    int cout, coutPad;             // real output channels; multiple of 16
    int relu;                      // max(0, .)
    const bf16* mask; int maskC;   // optional: times (mask[p][co] > 0)
    const bf16* resid; int residC; // optional: plus resid[p][co]
    bf16* dstB; float* dstF; int dstC;   // output (one of the two), channel stride dstC
    bf16* pool;                    // optional (bf16 output, no mask or residual): the 2x2 average
                                   // of dstB at (n, h/2, w/2), as launch_avgpool computes it
};
void launch_conv3x3(const ConvArgs& a, hipStream_t s);
// 0: each wave reads weights from L1/L2; 1: weight tiles staged in LDS per
// K step; 2 (default): input tiles with halo and weight slices staged in LDS.
void set_conv_variant(int v);
int conv_variant();

// dW[co][tap][ci] += sum_p g[p][co] * src(p + tap)[ci] for co < cout, src as
// in ConvArgs (c0 and c1 multiples of 32, w_ a multiple of 32); dB[co] +=
// sum_p g[p][co] when dB is non-null.  Partial sums of pixel chunks go to
// `part` and are added in chunk order (cout <= 128, c0 + c1 <= 256).
struct WgradArgs {
    const bf16* g; int gC;         // output gradient, gC channels (multiple of 16 >= cout)
    const bf16* src0; int c0; int up0;
    const bf16* src1; int c1;
    int n, h, w_;
    int cout, coutPad;             // weight rows; coutPad multiple of 16
    float* dW;                     // [coutPad][9][c0 + c1]
    float* dB;                     // [cout] or null
    float* part;                   // scratch for the chunk partial sums
};
// part holds kWgradScratchFloats: the chunk sums (at most kWgradPartFloats)
// followed by kWgradGroups group sums of every weight and bias gradient.
constexpr size_t kWgradPartFloats = (size_t)8 << 20;
constexpr int kWgradGroups = 16;
constexpr size_t kWgradScratchFloats = kWgradPartFloats + (size_t)kWgradGroups * (128 * 9 * 256 + 256);
void launch_wgrad3x3(const WgradArgs& a, hipStream_t s);

// 2x2 average pool: dst (n, h/2, w/2, c) from src (n, h, w, c).
void launch_avgpool(const bf16* src, bf16* dst, int n, int h, int w, int c, hipStream_t s);
// Gradient through the nearest 2x upsample of a concat input, times the ReLU
// mask of the upsampled tensor: out[q][c] = (sum of the 4 children of q in
// g[., c], c < ca) * (act[q][c] > 0); q at (n, h/2, w/2), g has gC channels.
void launch_upsample_bwd(const bf16* g, int gC, const bf16* act, bf16* out, int n, int h, int w, int ca,
                         hipStream_t s);
// Gradient of an encoder output that feeds both a pooling and a decoder
// concat: out[p][c] = (gp[p/2][c] / 4 + gcat[p][off + c]) * (act[p][c] > 0).
void launch_pool_bwd_add(const bf16* gp, const bf16* gcat, int gcatC, int off, const bf16* act, bf16* out, int n,
                         int h, int w, int c, hipStream_t s);

// Training loss over the 3 channels of every pixel (o, t in log(1 + RGB)):
// mode 0 = L1 in log space, mode 1 = relative MSE of the linear radiance with
// the prediction's own value (held constant) in the denominator.  Adds the
// mean loss to *lossAcc and writes go[p][c] = dLoss/do (c < 3), 0 up to goC.
void launch_loss(const float* o, const float* t, long long pixels, float invCount, int mode, bf16* go, int goC,
                 float* lossAcc, hipStream_t s);

// AdamW on one parameter block of `count` fp32 values.
struct AdamArgs {
    float* w; const float* g; float* m; float* v;
    long long count;
    float lr, beta1, beta2, eps, wd;
    float bc1, bc2;                // 1 - beta^t
    int decay;                     // apply weight decay (weights, not biases)
};
void launch_adamw(const AdamArgs& a, hipStream_t s);
// bf16 copies of a layer's fp32 weights: w16 = w, wt16[ci][8 - tap][co] = w[co][tap][ci]
// with wt16 of shape [cin][9][coutT] (coutT >= cout, zero-filled beyond cout).
void launch_weight_copies(const float* w, int coutPad, int cin, int coutT, bf16* w16, bf16* wt16, hipStream_t s);

// Render buffers -> network input: x[b][y][x][0..31] = {log(1+rgb'),
// albedo, normal, 0...} for y < h, x < w and 0 in the padding up to (hp, wp),
// rgb' = max(rgb, 0) after firefly suppression (kFireflyRatio),
// for b < flips (1, 2 or 4): image b is mirrored left-right when bit 0 of b is
// set, top-bottom when bit 1 is set (a flip ensemble at inference).
void launch_features(const float4* rgb, const float4* alb, const float4* nrm, int w, int h, int wp, int hp,
                     int flips, bf16* x, bf16* rgb4, hipStream_t s);
// The residual input of the output layer: channels 0-2 of the network input x
// (kInPad channels per pixel) and a zero, 4 channels per pixel (launch_features
// writes it along with x when rgb4 is non-null).
void launch_input_rgb(const bf16* x, long long pixels, bf16* rgb4, hipStream_t s);
// Network output (log space, fp32 [flips][hp][wp][3]) -> the mean over the
// mirrored-back images in log space -> linear float4 image and an optional
// gamma-packed RGBA8 buffer (the present step of the denoised image).
void launch_output(const float* o, int wp, int hp, int w, int h, int flips, float4* out, uint32_t* fb,
                   hipStream_t s);
// Targets for training: t[p][c] = log(1 + max(rgb, 0)).
void launch_log_target(const float4* rgb, long long pixels, float* t, hipStream_t s);
void launch_fill_f32(float* p, long long n, float v, hipStream_t s);
void launch_copy_bytes(void* dst, const void* src, long long bytes, hipStream_t s);

}  // namespace dn
}  // namespace bte
