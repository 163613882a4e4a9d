// bte_denoise_api.cpp — the learned denoiser behind the C ABI.
//
// Fills the reference's denoiser slot: cudaToggleDenoiser and
// denoiseAndPresent (src/host/denoiser.cu:72-116, called from
// src/render_kernel.cu:136-140 after each frame), whose closed OptiX model is
// replaced by a U-Net trained on this renderer's own low/high-spp pairs:
//
//   level 0 (H x W):     c0a 32->32, c0b 32->32 = s0            d0a [up(d1b), s0] 96->32, d0b 32->32, out 32->3
//   level 1 (H/2):       c1a 32->64, c1b 64->64 = s1            d1a [up(d2b), s1] 192->64, d1b 64->64
//   level 2 (H/4):       c2a 64->128, c2b 128->128 = s2         d2a [up(b3), s2] 256->128, d2b 128->128
//   level 3 (H/8):       c3a 128->128, c3b 128->128 = b3
//
// 3x3 convolutions with ReLU (none on `out`), 2x2 average pooling down,
// nearest 2x upsampling up, and a residual: the output (log(1 + RGB)) is the
// input's log channels plus the network's last convolution.  Input channels:
// log(1 + max(RGB, 0)) of the accumulation, albedo and normal AOVs.  Training
// minimises the mean L1 distance in that log space (or, selectable, the
// relative MSE of the linear radiance) with AdamW.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bte_denoise.h"
#include "bte_internal.h"

namespace bte {
namespace {

using dn::bf16;

struct LayerDef {
    int cinReal, cin, cout, relu;
};
constexpr LayerDef kLayers[] = {
    {9, 32, 32, 1},    {32, 32, 32, 1},   {32, 32, 64, 1},   {64, 64, 64, 1},    {64, 64, 128, 1},
    {128, 128, 128, 1}, {128, 128, 128, 1}, {128, 128, 128, 1}, {256, 256, 128, 1}, {128, 128, 128, 1},
    {192, 192, 64, 1}, {64, 64, 64, 1},   {96, 96, 32, 1},   {32, 32, 32, 1},    {32, 32, 3, 0},
};
enum { C0A, C0B, C1A, C1B, C2A, C2B, C3A, C3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT, kNumLayers };

struct Layer {
    int cinReal, cin, cout, coutPad, coutT, relu;
    size_t wOff, bOff, w16Off, wt16Off;   // offsets in the fp32 / bf16 parameter buffers
};

// Activation and gradient tensors (bf16 NHWC) with their level and channels.
enum Tid {
    T_X, T_E0A, T_S0, T_P0, T_E1A, T_S1, T_P1, T_E2A, T_S2, T_P2, T_E3A, T_B3, T_D2A, T_D2B, T_D1A, T_D1B, T_D0A,
    T_D0B, kForwardTensors,
    T_GO = kForwardTensors, T_GD0B, T_GD0A, T_GCAT0, T_GD1B, T_GD1A, T_GCAT1, T_GD2B, T_GD2A, T_GCAT2, T_GB3, T_GE3A,
    T_GP2, T_GS2, T_GE2A, T_GP1, T_GS1, T_GE1A, T_GP0, T_GS0, T_GE0A, kTensors
};
struct TDesc {
    int level, c;
};
constexpr TDesc kT[kTensors] = {
    {0, 32}, {0, 32}, {0, 32}, {1, 32}, {1, 64}, {1, 64}, {2, 64}, {2, 128}, {2, 128}, {3, 128}, {3, 128},
    {3, 128}, {2, 128}, {2, 128}, {1, 64}, {1, 64}, {0, 32}, {0, 32},
    // gradients
    {0, 32}, {0, 32}, {0, 32}, {0, 96}, {1, 64}, {1, 64}, {1, 192}, {2, 128}, {2, 128}, {2, 256}, {3, 128},
    {3, 128}, {3, 128}, {2, 128}, {2, 128}, {2, 64}, {1, 64}, {1, 64}, {1, 32}, {0, 32}, {0, 32},
};
constexpr int kOutputTensor = kTensors;   // id of the fp32 network output in the tensor API

struct Net {
    bool ready = false;
    Layer L[kNumLayers];
    size_t nW = 0, nW16 = 0, nWt16 = 0;   // fp32 parameter count (weights + biases), bf16 sizes
    size_t nWeightsOnly = 0;              // fp32 weights before the biases
    DBuf w, g, m, v, w16, wt16;
    long long step = 0;
    // workspace
    int n = 0, h = 0, wdt = 0;
    bool train = false;
    DBuf t[kTensors];
    DBuf o, target, loss;
    DBuf wgPart;                 // weight-gradient chunk sums
    DBuf rgb4;                   // channels 0-2 of the input, 4 per pixel: the output layer's residual
    // dataset of training pairs
    int slots = 0, dsW = 0, dsH = 0;
    DBuf dsX, dsT;
    std::vector<char> dsHave;   // bit 0 input captured, bit 1 target captured
    // inference
    DBuf denoised;
    int outW = 0, outH = 0;
    bool enabled = false;
    int lossMode = 0;   // 0: L1 in log space, 1: relative MSE (linear)
    int flips = 4;      // inference: mean over 1, 2 or 4 mirrored copies (log space)
};

Net& net() {
    static Net n;
    return n;
}

void buildLayout(Net& N) {
    size_t wOff = 0, w16 = 0, wt16 = 0;
    for (int i = 0; i < kNumLayers; ++i) {
        Layer& l = N.L[i];
        l.cinReal = kLayers[i].cinReal;
        l.cin = kLayers[i].cin;
        l.cout = kLayers[i].cout;
        l.relu = kLayers[i].relu;
        l.coutPad = (l.cout + 15) / 16 * 16;
        l.coutT = (l.cout + 31) / 32 * 32;   // channels of this layer's output gradient
        l.wOff = wOff;
        wOff += (size_t)l.coutPad * 9 * l.cin;
        l.w16Off = w16;
        w16 += (size_t)l.coutPad * 9 * l.cin;
        l.wt16Off = wt16;
        wt16 += (size_t)l.cin * 9 * l.coutT;
    }
    N.nWeightsOnly = wOff;
    for (int i = 0; i < kNumLayers; ++i) {
        N.L[i].bOff = wOff;
        wOff += (size_t)N.L[i].coutPad;
    }
    N.nW = wOff;
    N.nW16 = w16;
    N.nWt16 = wt16;
}

long long canonicalCount() {
    long long n = 0;
    for (const LayerDef& d : kLayers) n += (long long)d.cout * 9 * d.cinReal + d.cout;
    return n;
}

// canonical host layout per layer: W[cout][3][3][cinReal], then b[cout]
void canonicalToPadded(const Net& N, const float* src, std::vector<float>& dst) {
    dst.assign(N.nW, 0.0f);
    size_t k = 0;
    for (int i = 0; i < kNumLayers; ++i) {
        const Layer& l = N.L[i];
        for (int co = 0; co < l.cout; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < l.cinReal; ++ci) dst[l.wOff + ((size_t)co * 9 + t) * l.cin + ci] = src[k++];
        for (int co = 0; co < l.cout; ++co) dst[l.bOff + co] = src[k++];
    }
}

void paddedToCanonical(const Net& N, const std::vector<float>& src, float* dst) {
    size_t k = 0;
    for (int i = 0; i < kNumLayers; ++i) {
        const Layer& l = N.L[i];
        for (int co = 0; co < l.cout; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < l.cinReal; ++ci) dst[k++] = src[l.wOff + ((size_t)co * 9 + t) * l.cin + ci];
        for (int co = 0; co < l.cout; ++co) dst[k++] = src[l.bOff + co];
    }
}

int ensureParams(Net& N, hipStream_t s) {
    if (N.ready) return BTE_OK;
    buildLayout(N);
    CHECK_RC(N.w.ensure(N.nW * sizeof(float)));
    CHECK_RC(N.g.ensure(N.nW * sizeof(float)));
    CHECK_RC(N.m.ensure(N.nW * sizeof(float)));
    CHECK_RC(N.v.ensure(N.nW * sizeof(float)));
    CHECK_RC(N.w16.ensure(N.nW16 * sizeof(bf16)));
    CHECK_RC(N.wt16.ensure(N.nWt16 * sizeof(bf16)));
    HIPCHK(hipMemsetAsync(N.wt16.p, 0, N.nWt16 * sizeof(bf16), s));
    HIPCHK(hipStreamSynchronize(s));
    N.ready = true;
    return BTE_OK;
}

void refreshCopies(Net& N, hipStream_t s) {
    for (int i = 0; i < kNumLayers; ++i) {
        const Layer& l = N.L[i];
        dn::launch_weight_copies(N.w.as<float>() + l.wOff, l.coutPad, l.cin, l.coutT, N.w16.as<bf16>() + l.w16Off,
                                 N.wt16.as<bf16>() + l.wt16Off, s);
    }
}

// Uploads padded fp32 parameters, clears the optimiser state.
int uploadParams(Net& N, const std::vector<float>& padded, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(N.w.p, padded.data(), N.nW * sizeof(float), hipMemcpyHostToDevice, s));
    dn::launch_fill_f32(N.m.as<float>(), (long long)N.nW, 0.0f, s);
    dn::launch_fill_f32(N.v.as<float>(), (long long)N.nW, 0.0f, s);
    refreshCopies(N, s);
    HIPCHK(hipStreamSynchronize(s));
    N.step = 0;
    return BTE_OK;
}

int ensureWorkspace(Net& N, int n, int h, int w, bool train) {
    if (h % 8 || w % 8 || n <= 0 || h <= 0 || w <= 0)
        return fail(BTE_ERR_INVALID, "denoiser tensors need positive sizes with H, W multiples of 8 (got %dx%dx%d)", n, h, w);
    const int count = train ? (int)kTensors : (int)kForwardTensors;
    for (int i = 0; i < count; ++i) {
        const size_t px = (size_t)n * (size_t)(h >> kT[i].level) * (size_t)(w >> kT[i].level);
        CHECK_RC(N.t[i].ensure(px * kT[i].c * sizeof(bf16)));
    }
    const size_t P = (size_t)n * h * w;
    CHECK_RC(N.o.ensure(P * 3 * sizeof(float)));
    CHECK_RC(N.rgb4.ensure(P * 4 * sizeof(bf16)));
    if (train) {
        CHECK_RC(N.target.ensure(P * 3 * sizeof(float)));
        CHECK_RC(N.loss.ensure(sizeof(float)));
        CHECK_RC(N.wgPart.ensure(dn::kWgradScratchFloats * sizeof(float)));
    }
    N.n = n;
    N.h = h;
    N.wdt = w;
    N.train = N.train || train;
    return BTE_OK;
}

bf16* T(Net& N, int id) { return N.t[id].as<bf16>(); }

struct Src {
    const bf16* s0; int c0; int up0;
    const bf16* s1; int c1;
};

void conv(Net& N, int layer, const Src& in, int level, bf16* dstB, float* dstF, const bf16* resid, hipStream_t s,
          bf16* pool = nullptr) {
    const Layer& l = N.L[layer];
    dn::ConvArgs a{};
    a.src0 = in.s0; a.c0 = in.c0; a.up0 = in.up0;
    a.src1 = in.s1; a.c1 = in.c1;
    a.w = N.w16.as<bf16>() + l.w16Off;
    a.bias = N.w.as<float>() + l.bOff;
    a.n = N.n; a.h = N.h >> level; a.w_ = N.wdt >> level;
    a.cout = l.cout; a.coutPad = l.coutPad;
    a.relu = l.relu;
    a.resid = resid; a.residC = 4;
    a.dstB = dstB; a.dstF = dstF; a.dstC = dstF ? dn::kOutChannels : l.cout;
    a.pool = pool;
    dn::launch_conv3x3(a, s);
}

// gradient w.r.t. the layer's input (times the ReLU mask of that input when given)
void dgrad(Net& N, int layer, const bf16* g, int level, const bf16* mask, bf16* dst, hipStream_t s) {
    const Layer& l = N.L[layer];
    dn::ConvArgs a{};
    a.src0 = g; a.c0 = l.coutT; a.up0 = 0;
    a.src1 = nullptr; a.c1 = 0;
    a.w = N.wt16.as<bf16>() + l.wt16Off;
    a.bias = nullptr;
    a.n = N.n; a.h = N.h >> level; a.w_ = N.wdt >> level;
    a.cout = l.cin; a.coutPad = l.cin;
    a.relu = 0;
    a.mask = mask; a.maskC = l.cin;
    a.dstB = dst; a.dstF = nullptr; a.dstC = l.cin;
    dn::launch_conv3x3(a, s);
}

void wgrad(Net& N, int layer, const bf16* g, const Src& in, int level, hipStream_t s) {
    const Layer& l = N.L[layer];
    dn::WgradArgs a{};
    a.g = g; a.gC = l.coutT;
    a.src0 = in.s0; a.c0 = in.c0; a.up0 = in.up0;
    a.src1 = in.s1; a.c1 = in.c1;
    a.n = N.n; a.h = N.h >> level; a.w_ = N.wdt >> level;
    a.cout = l.cout; a.coutPad = l.coutPad;
    a.dW = N.g.as<float>() + l.wOff;
    a.dB = N.g.as<float>() + l.bOff;
    a.part = N.wgPart.as<float>();
    dn::launch_wgrad3x3(a, s);
}

Src one(const bf16* s0, int c0) { return Src{s0, c0, 0, nullptr, 0}; }
Src cat(const bf16* up, int cu, const bf16* skip, int cs) { return Src{up, cu, 1, skip, cs}; }

void forward(Net& N, hipStream_t s) {
    conv(N, C0A, one(T(N, T_X), 32), 0, T(N, T_E0A), nullptr, nullptr, s);
    conv(N, C0B, one(T(N, T_E0A), 32), 0, T(N, T_S0), nullptr, nullptr, s, T(N, T_P0));
    conv(N, C1A, one(T(N, T_P0), 32), 1, T(N, T_E1A), nullptr, nullptr, s);
    conv(N, C1B, one(T(N, T_E1A), 64), 1, T(N, T_S1), nullptr, nullptr, s, T(N, T_P1));
    conv(N, C2A, one(T(N, T_P1), 64), 2, T(N, T_E2A), nullptr, nullptr, s);
    conv(N, C2B, one(T(N, T_E2A), 128), 2, T(N, T_S2), nullptr, nullptr, s, T(N, T_P2));
    conv(N, C3A, one(T(N, T_P2), 128), 3, T(N, T_E3A), nullptr, nullptr, s);
    conv(N, C3B, one(T(N, T_E3A), 128), 3, T(N, T_B3), nullptr, nullptr, s);
    conv(N, D2A, cat(T(N, T_B3), 128, T(N, T_S2), 128), 2, T(N, T_D2A), nullptr, nullptr, s);
    conv(N, D2B, one(T(N, T_D2A), 128), 2, T(N, T_D2B), nullptr, nullptr, s);
    conv(N, D1A, cat(T(N, T_D2B), 128, T(N, T_S1), 64), 1, T(N, T_D1A), nullptr, nullptr, s);
    conv(N, D1B, one(T(N, T_D1A), 64), 1, T(N, T_D1B), nullptr, nullptr, s);
    conv(N, D0A, cat(T(N, T_D1B), 64, T(N, T_S0), 32), 0, T(N, T_D0A), nullptr, nullptr, s);
    conv(N, D0B, one(T(N, T_D0A), 32), 0, T(N, T_D0B), nullptr, nullptr, s);
    conv(N, OUT, one(T(N, T_D0B), 32), 0, nullptr, N.o.as<float>(), N.rgb4.as<bf16>(), s);
}

// Gradients of the mean L1 loss into N.g (cleared first); loss into N.loss.
void backward(Net& N, hipStream_t s) {
    const int n = N.n, h = N.h, w = N.wdt;
    const long long P = (long long)n * h * w;
    dn::launch_fill_f32(N.g.as<float>(), (long long)N.nW, 0.0f, s);
    dn::launch_fill_f32(N.loss.as<float>(), 1, 0.0f, s);
    dn::launch_loss(N.o.as<float>(), N.target.as<float>(), P, 1.0f / (float)(P * 3), N.lossMode, T(N, T_GO), 32,
                    N.loss.as<float>(), s);
    // decoder, level 0
    wgrad(N, OUT, T(N, T_GO), one(T(N, T_D0B), 32), 0, s);
    dgrad(N, OUT, T(N, T_GO), 0, T(N, T_D0B), T(N, T_GD0B), s);
    wgrad(N, D0B, T(N, T_GD0B), one(T(N, T_D0A), 32), 0, s);
    dgrad(N, D0B, T(N, T_GD0B), 0, T(N, T_D0A), T(N, T_GD0A), s);
    wgrad(N, D0A, T(N, T_GD0A), cat(T(N, T_D1B), 64, T(N, T_S0), 32), 0, s);
    dgrad(N, D0A, T(N, T_GD0A), 0, nullptr, T(N, T_GCAT0), s);
    dn::launch_upsample_bwd(T(N, T_GCAT0), 96, T(N, T_D1B), T(N, T_GD1B), n, h, w, 64, s);
    // level 1
    wgrad(N, D1B, T(N, T_GD1B), one(T(N, T_D1A), 64), 1, s);
    dgrad(N, D1B, T(N, T_GD1B), 1, T(N, T_D1A), T(N, T_GD1A), s);
    wgrad(N, D1A, T(N, T_GD1A), cat(T(N, T_D2B), 128, T(N, T_S1), 64), 1, s);
    dgrad(N, D1A, T(N, T_GD1A), 1, nullptr, T(N, T_GCAT1), s);
    dn::launch_upsample_bwd(T(N, T_GCAT1), 192, T(N, T_D2B), T(N, T_GD2B), n, h / 2, w / 2, 128, s);
    // level 2
    wgrad(N, D2B, T(N, T_GD2B), one(T(N, T_D2A), 128), 2, s);
    dgrad(N, D2B, T(N, T_GD2B), 2, T(N, T_D2A), T(N, T_GD2A), s);
    wgrad(N, D2A, T(N, T_GD2A), cat(T(N, T_B3), 128, T(N, T_S2), 128), 2, s);
    dgrad(N, D2A, T(N, T_GD2A), 2, nullptr, T(N, T_GCAT2), s);
    dn::launch_upsample_bwd(T(N, T_GCAT2), 256, T(N, T_B3), T(N, T_GB3), n, h / 4, w / 4, 128, s);
    // bottleneck, level 3
    wgrad(N, C3B, T(N, T_GB3), one(T(N, T_E3A), 128), 3, s);
    dgrad(N, C3B, T(N, T_GB3), 3, T(N, T_E3A), T(N, T_GE3A), s);
    wgrad(N, C3A, T(N, T_GE3A), one(T(N, T_P2), 128), 3, s);
    dgrad(N, C3A, T(N, T_GE3A), 3, nullptr, T(N, T_GP2), s);
    // encoder
    dn::launch_pool_bwd_add(T(N, T_GP2), T(N, T_GCAT2), 256, 128, T(N, T_S2), T(N, T_GS2), n, h / 4, w / 4, 128, s);
    wgrad(N, C2B, T(N, T_GS2), one(T(N, T_E2A), 128), 2, s);
    dgrad(N, C2B, T(N, T_GS2), 2, T(N, T_E2A), T(N, T_GE2A), s);
    wgrad(N, C2A, T(N, T_GE2A), one(T(N, T_P1), 64), 2, s);
    dgrad(N, C2A, T(N, T_GE2A), 2, nullptr, T(N, T_GP1), s);
    dn::launch_pool_bwd_add(T(N, T_GP1), T(N, T_GCAT1), 192, 128, T(N, T_S1), T(N, T_GS1), n, h / 2, w / 2, 64, s);
    wgrad(N, C1B, T(N, T_GS1), one(T(N, T_E1A), 64), 1, s);
    dgrad(N, C1B, T(N, T_GS1), 1, T(N, T_E1A), T(N, T_GE1A), s);
    wgrad(N, C1A, T(N, T_GE1A), one(T(N, T_P0), 32), 1, s);
    dgrad(N, C1A, T(N, T_GE1A), 1, nullptr, T(N, T_GP0), s);
    dn::launch_pool_bwd_add(T(N, T_GP0), T(N, T_GCAT0), 96, 64, T(N, T_S0), T(N, T_GS0), n, h, w, 32, s);
    wgrad(N, C0B, T(N, T_GS0), one(T(N, T_E0A), 32), 0, s);
    dgrad(N, C0B, T(N, T_GS0), 0, T(N, T_E0A), T(N, T_GE0A), s);
    wgrad(N, C0A, T(N, T_GE0A), one(T(N, T_X), 32), 0, s);
}

void adamStep(Net& N, float lr, float wd, hipStream_t s) {
    N.step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    dn::AdamArgs a{};
    a.w = N.w.as<float>(); a.g = N.g.as<float>(); a.m = N.m.as<float>(); a.v = N.v.as<float>();
    a.lr = lr; a.beta1 = b1; a.beta2 = b2; a.eps = 1e-8f; a.wd = wd;
    a.bc1 = 1.0f - (float)std::pow((double)b1, (double)N.step);
    a.bc2 = 1.0f - (float)std::pow((double)b2, (double)N.step);
    a.count = (long long)N.nWeightsOnly;
    a.decay = 1;
    dn::launch_adamw(a, s);
    a.w += N.nWeightsOnly; a.g += N.nWeightsOnly; a.m += N.nWeightsOnly; a.v += N.nWeightsOnly;
    a.count = (long long)(N.nW - N.nWeightsOnly);
    a.decay = 0;
    dn::launch_adamw(a, s);
    refreshCopies(N, s);
}

// float -> bf16, round to nearest even (the conversion the kernels use)
uint16_t toBf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float fromBf16(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

int uploadInput(Net& N, const float* x9, hipStream_t s) {
    const size_t P = (size_t)N.n * N.h * N.wdt;
    std::vector<uint16_t> buf(P * dn::kInPad, 0);
    for (size_t p = 0; p < P; ++p)
        for (int c = 0; c < dn::kInChannels; ++c) buf[p * dn::kInPad + c] = toBf16(x9[p * dn::kInChannels + c]);
    HIPCHK(hipMemcpyAsync(N.t[T_X].p, buf.data(), buf.size() * 2, hipMemcpyHostToDevice, s));
    dn::launch_input_rgb(T(N, T_X), (long long)P, N.rgb4.as<bf16>(), s);
    HIPCHK(hipStreamSynchronize(s));
    return BTE_OK;
}

#define NET_READY()                                                                   \
    do {                                                                              \
        CHECK_RC(ensure_device());                                                    \
        if (!net().ready) return fail(BTE_ERR_STATE, "denoiser weights not loaded"); \
    } while (0)

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

bool denoiser_active() { return net().enabled && net().ready; }

int denoise_view(const RenderView& v, float4* out, bool captured) {
    Net& N = net();
    if (!N.ready) return fail(BTE_ERR_STATE, "denoiser weights not loaded");
    if (!v.aovAlb) return fail(BTE_ERR_STATE, "the denoiser needs the albedo/normal AOVs (bte_enable_aov)");
    const int hp = (v.height + 7) / 8 * 8, wp = (v.width + 7) / 8 * 8;
    // the flip ensemble runs as one batch of N.flips mirrored images
    CHECK_RC(ensureWorkspace(N, N.flips, hp, wp, false));
    auto emit = [&]() {
        dn::launch_features(v.accum, v.aovAlb, v.aovNrm, v.width, v.height, wp, hp, N.flips, T(N, T_X),
                            N.rgb4.as<bf16>(), v.stream);
        forward(N, v.stream);
        dn::launch_output(N.o.as<float>(), wp, hp, v.width, v.height, N.flips, out, v.fb, v.stream);
    };
    if (!captured || v.stream != main_stream()) {
        emit();
        return BTE_OK;
    }
    // a graph of the ~18 launches: every buffer and size they use is in the key
    // (weights are read from their buffers at replay, so new values need no
    // new graph)
    std::vector<unsigned char> key;
    auto add = [&](const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        key.insert(key.end(), b, b + n);
    };
    const long long sizes[] = {0x64656e6f697365ll, N.flips, wp, hp, v.width, v.height, dn::conv_variant()};
    add(sizes, sizeof sizes);
    const void* ptrs[] = {v.accum, v.aovAlb, v.aovNrm, out, v.fb, N.o.p, N.rgb4.p, N.w16.p, N.w.p};
    add(ptrs, sizeof ptrs);
    for (const DBuf& b : N.t) add(&b.p, sizeof b.p);
    return run_captured(key, emit);
}

int denoiser_present() {
    Net& N = net();
    RenderView v;
    CHECK_RC(render_view(v));
    CHECK_RC(N.denoised.ensure((size_t)v.width * v.height * sizeof(float4)));
    CHECK_RC(denoise_view(v, N.denoised.as<float4>(), true));
    HIPCHK(hipStreamSynchronize(v.stream));
    HIPCHK(hipGetLastError());
    N.outW = v.width;
    N.outH = v.height;
    return BTE_OK;
}

// bte_cleanup: frees the per-resolution state and switches the denoiser off
// (src/host/denoiser.cu:121-129); the model and the training pairs stay, so
// training can move from scene to scene (bte_denoiser_release frees them).
void denoiser_release() {
    Net& N = net();
    for (DBuf& b : N.t) b.release();
    N.o.release(); N.target.release(); N.loss.release(); N.wgPart.release(); N.rgb4.release();
    N.denoised.release();
    N.outW = N.outH = 0;
    N.enabled = false;
    N.n = N.h = N.wdt = 0;
    N.train = false;
}

}  // namespace bte

using namespace bte;

extern "C" {

int bte_toggle_denoiser(void) {
    Net& N = net();
    if (!N.ready) return fail(BTE_ERR_STATE, "denoiser weights not loaded (bte_denoiser_init / set_weights)");
    N.enabled = !N.enabled;
    return N.enabled ? 1 : 0;
}

int bte_denoiser_release(void) {
    Net& N = net();
    denoiser_release();
    N.dsX.release(); N.dsT.release();
    N.slots = 0;
    N.dsHave.clear();
    N.w.release(); N.g.release(); N.m.release(); N.v.release(); N.w16.release(); N.wt16.release();
    N.ready = false;
    N.step = 0;
    return BTE_OK;
}

int bte_denoise(void) {
    NET_READY();
    return denoiser_present();
}

int bte_read_denoised(float* hostRGB, size_t count) {
    CHECK_RC(ensure_device());
    Net& N = net();
    const size_t P = (size_t)N.outW * N.outH;
    if (!N.denoised.p || P == 0) return fail(BTE_ERR_STATE, "nothing denoised yet");
    if (!hostRGB || count < 3 * P) return fail(BTE_ERR_INVALID, "denoised copy needs %zu floats", 3 * P);
    std::vector<float> tmp(4 * P);
    HIPCHK(hipMemcpy(tmp.data(), N.denoised.p, P * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

long long bte_denoiser_param_count(void) { return canonicalCount(); }

int bte_denoiser_init(unsigned seed) {
    CHECK_RC(ensure_device());
    hipStream_t s = main_stream();
    Net& N = net();
    buildLayout(N);
    CHECK_RC(ensureParams(N, s));
    // He-normal weights (std sqrt(2 / fan_in)), zero biases; the last layer
    // starts 100x smaller so the residual output begins near the input.
    std::vector<float> canon((size_t)canonicalCount(), 0.0f);
    uint64_t st = 0x5eedull ^ ((uint64_t)seed << 20);
    size_t k = 0;
    for (int i = 0; i < kNumLayers; ++i) {
        const LayerDef& d = kLayers[i];
        const double sd = std::sqrt(2.0 / (9.0 * d.cinReal)) * (i == OUT ? 0.01 : 1.0);
        const size_t nw = (size_t)d.cout * 9 * d.cinReal;
        for (size_t j = 0; j < nw; ++j) {
            const double u1 = ((double)(splitmix(st) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
            const double u2 = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            canon[k++] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        }
        k += (size_t)d.cout;   // biases stay 0
    }
    std::vector<float> padded;
    canonicalToPadded(N, canon.data(), padded);
    return uploadParams(N, padded, s);
}

int bte_denoiser_set_weights(const float* params, long long count) {
    CHECK_RC(ensure_device());
    if (!params || count != canonicalCount())
        return fail(BTE_ERR_INVALID, "denoiser weights: expected %lld floats", canonicalCount());
    Net& N = net();
    hipStream_t s = main_stream();
    buildLayout(N);
    CHECK_RC(ensureParams(N, s));
    std::vector<float> padded;
    canonicalToPadded(N, params, padded);
    return uploadParams(N, padded, s);
}

// Weights file: a NumPy .npy array of little-endian float32 in canonical order.
int bte_denoiser_load_file(const char* path) {
    if (!path) return fail(BTE_ERR_INVALID, "null path");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(BTE_ERR_INVALID, "cannot open %s", path);
    unsigned char head[10];
    std::vector<float> w;
    int rc = BTE_OK;
    if (std::fread(head, 1, 10, fp) != 10 || std::memcmp(head, "\x93NUMPY", 6) != 0) {
        rc = fail(BTE_ERR_INVALID, "%s is not a .npy file", path);
    } else {
        size_t hlen = 0;
        if (head[6] == 1) {
            hlen = (size_t)head[8] | ((size_t)head[9] << 8);
        } else {
            unsigned char ext[2];
            if (std::fread(ext, 1, 2, fp) != 2) hlen = 0;
            else hlen = (size_t)head[8] | ((size_t)head[9] << 8) | ((size_t)ext[0] << 16) | ((size_t)ext[1] << 24);
        }
        std::string hdr(hlen, '\0');
        if (hlen == 0 || std::fread(&hdr[0], 1, hlen, fp) != hlen || hdr.find("'<f4'") == std::string::npos ||
            hdr.find("'fortran_order': False") == std::string::npos) {
            rc = fail(BTE_ERR_INVALID, "%s: expected a C-ordered little-endian float32 array", path);
        } else {
            const long long n = canonicalCount();
            w.resize((size_t)n);
            if (std::fread(w.data(), sizeof(float), (size_t)n, fp) != (size_t)n || std::fgetc(fp) != EOF)
                rc = fail(BTE_ERR_INVALID, "%s: expected exactly %lld floats", path, n);
        }
    }
    std::fclose(fp);
    if (rc != BTE_OK) return rc;
    return bte_denoiser_set_weights(w.data(), (long long)w.size());
}

static int readCanonical(const DBuf& b, float* out, long long count) {
    NET_READY();
    Net& N = net();
    if (!out || count != canonicalCount())
        return fail(BTE_ERR_INVALID, "denoiser parameters: expected %lld floats", canonicalCount());
    std::vector<float> padded(N.nW);
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(padded.data(), b.p, N.nW * sizeof(float), hipMemcpyDeviceToHost));
    paddedToCanonical(N, padded, out);
    return BTE_OK;
}

int bte_denoiser_get_weights(float* params, long long count) { return readCanonical(net().w, params, count); }
int bte_denoiser_get_grads(float* grads, long long count) { return readCanonical(net().g, grads, count); }

int bte_denoiser_forward_host(const float* x9, int n, int h, int w, float* out3) {
    NET_READY();
    if (!x9 || !out3) return fail(BTE_ERR_INVALID, "null buffer");
    Net& N = net();
    CHECK_RC(ensureWorkspace(N, n, h, w, false));
    hipStream_t s = main_stream();
    CHECK_RC(uploadInput(N, x9, s));
    forward(N, s);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(out3, N.o.p, (size_t)n * h * w * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_denoiser_train_step_host(const float* x9, const float* target3, int n, int h, int w, float lr, float wd,
                                 float* loss) {
    NET_READY();
    if (!x9 || !target3) return fail(BTE_ERR_INVALID, "null buffer");
    if (w % 256) return fail(BTE_ERR_INVALID, "training tiles need a width that is a multiple of 256 (got %d)", w);
    Net& N = net();
    CHECK_RC(ensureWorkspace(N, n, h, w, true));
    hipStream_t s = main_stream();
    CHECK_RC(uploadInput(N, x9, s));
    HIPCHK(hipMemcpyAsync(N.target.p, target3, (size_t)n * h * w * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    forward(N, s);
    backward(N, s);
    if (lr > 0.0f) adamStep(N, lr, wd, s);
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (loss) HIPCHK(hipMemcpy(loss, N.loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_denoiser_tensor_shape(int id, int* shape4) {
    Net& N = net();
    if (!shape4) return fail(BTE_ERR_INVALID, "null shape");
    if (id == kOutputTensor) {
        shape4[0] = N.n; shape4[1] = N.h; shape4[2] = N.wdt; shape4[3] = 3;
        return BTE_OK;
    }
    if (id < 0 || id >= kTensors) return fail(BTE_ERR_INVALID, "tensor id %d out of range", id);
    shape4[0] = N.n;
    shape4[1] = N.h >> kT[id].level;
    shape4[2] = N.wdt >> kT[id].level;
    shape4[3] = kT[id].c;
    return BTE_OK;
}

int bte_denoiser_read_tensor(int id, float* out, long long count) {
    NET_READY();
    Net& N = net();
    int sh[4];
    CHECK_RC(bte_denoiser_tensor_shape(id, sh));
    const long long total = (long long)sh[0] * sh[1] * sh[2] * sh[3];
    if (!out || count < total) return fail(BTE_ERR_INVALID, "tensor %d needs %lld floats", id, total);
    HIPCHK(hipDeviceSynchronize());
    if (id == kOutputTensor) {
        HIPCHK(hipMemcpy(out, N.o.p, (size_t)total * sizeof(float), hipMemcpyDeviceToHost));
        return BTE_OK;
    }
    if (!N.t[id].p || N.t[id].bytes < (size_t)total * 2) return fail(BTE_ERR_STATE, "tensor %d not computed", id);
    std::vector<uint16_t> tmp((size_t)total);
    HIPCHK(hipMemcpy(tmp.data(), N.t[id].p, (size_t)total * 2, hipMemcpyDeviceToHost));
    for (long long i = 0; i < total; ++i) out[i] = fromBf16(tmp[(size_t)i]);
    return BTE_OK;
}

int bte_denoise_image(const float* rgb, const float* albedo, const float* normal, int width, int height,
                      float* out) {
    CHECK_RC(ensure_device());
    if (!rgb || !albedo || !normal || !out || width <= 0 || height <= 0)
        return fail(BTE_ERR_INVALID, "bte_denoise_image: null buffer or empty size");
    for (const void* p : {(const void*)rgb, (const void*)albedo, (const void*)normal, (const void*)out})
        if (reinterpret_cast<uintptr_t>(p) & 15u)
            return fail(BTE_ERR_INVALID, "bte_denoise_image: buffers must be 16-byte aligned float4 arrays");
    RenderView v;
    v.stream = main_stream();
    v.width = width;
    v.height = height;
    v.accum = reinterpret_cast<const float4*>(rgb);
    v.aovAlb = reinterpret_cast<const float4*>(albedo);
    v.aovNrm = reinterpret_cast<const float4*>(normal);
    CHECK_RC(denoise_view(v, reinterpret_cast<float4*>(out)));   // caller buffers vary: no graph per call
    HIPCHK(hipStreamSynchronize(v.stream));
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

int bte_denoiser_set_conv_kernel(int variant) {
    if (variant < 0 || variant > 2) return fail(BTE_ERR_INVALID, "conv kernel variant must be 0, 1 or 2 (got %d)", variant);
    dn::set_conv_variant(variant);
    return BTE_OK;
}

int bte_denoiser_set_flips(int flips) {
    if (flips != 1 && flips != 2 && flips != 4) return fail(BTE_ERR_INVALID, "flips must be 1, 2 or 4 (got %d)", flips);
    net().flips = flips;
    return BTE_OK;
}

int bte_denoiser_get_flips(void) { return net().flips; }

int bte_denoiser_set_loss(int mode) {
    if (mode != 0 && mode != 1) return fail(BTE_ERR_INVALID, "loss mode must be 0 (L1, log space) or 1 (relMSE)");
    net().lossMode = mode;
    return BTE_OK;
}

int bte_denoiser_dataset_reserve(int slots, int width, int height) {
    NET_READY();
    if (slots <= 0 || width <= 0 || height <= 0 || width % 256 || height % 8)
        return fail(BTE_ERR_INVALID, "dataset tiles need width a multiple of 256 and height of 8");
    Net& N = net();
    const size_t P = (size_t)width * height;
    CHECK_RC(N.dsX.ensure((size_t)slots * P * dn::kInPad * sizeof(bf16)));
    CHECK_RC(N.dsT.ensure((size_t)slots * P * 3 * sizeof(float)));
    N.slots = slots;
    N.dsW = width;
    N.dsH = height;
    N.dsHave.assign((size_t)slots, 0);
    return BTE_OK;
}

int bte_denoiser_capture(int slot, int which) {
    NET_READY();
    Net& N = net();
    if (slot < 0 || slot >= N.slots) return fail(BTE_ERR_INVALID, "slot %d outside the reserved %d", slot, N.slots);
    if (which != 0 && which != 1) return fail(BTE_ERR_INVALID, "which must be 0 (input) or 1 (target)");
    RenderView v;
    CHECK_RC(render_view(v));
    if (v.width != N.dsW || v.height != N.dsH)
        return fail(BTE_ERR_INVALID, "render is %dx%d, dataset tiles are %dx%d", v.width, v.height, N.dsW, N.dsH);
    const size_t P = (size_t)N.dsW * N.dsH;
    if (which == 0) {
        if (!v.aovAlb) return fail(BTE_ERR_STATE, "capture of inputs needs the AOVs (bte_enable_aov)");
        dn::launch_features(v.accum, v.aovAlb, v.aovNrm, v.width, v.height, v.width, v.height, 1,
                            N.dsX.as<bf16>() + (size_t)slot * P * dn::kInPad, nullptr, v.stream);
    } else {
        dn::launch_log_target(v.accum, (long long)P, N.dsT.as<float>() + (size_t)slot * P * 3, v.stream);
    }
    HIPCHK(hipStreamSynchronize(v.stream));
    N.dsHave[(size_t)slot] |= (char)(1 << which);
    return BTE_OK;
}

int bte_denoiser_train(int steps, int batch, float lr, float weightDecay, unsigned seed, float* losses) {
    NET_READY();
    Net& N = net();
    if (steps <= 0 || batch <= 0) return fail(BTE_ERR_INVALID, "steps and batch must be positive");
    std::vector<int> ready;
    for (int i = 0; i < N.slots; ++i)
        if (N.dsHave[(size_t)i] == 3) ready.push_back(i);
    if (ready.empty()) return fail(BTE_ERR_STATE, "no complete training pair captured");
    CHECK_RC(ensureWorkspace(N, batch, N.dsH, N.dsW, true));
    hipStream_t s = main_stream();
    DBuf lossHist;
    CHECK_RC(lossHist.ensure((size_t)steps * sizeof(float)));
    const size_t P = (size_t)N.dsW * N.dsH;
    uint64_t st = 0x7a11ull ^ ((uint64_t)seed << 17);
    for (int it = 0; it < steps; ++it) {
        for (int b = 0; b < batch; ++b) {
            const int slot = ready[(size_t)(splitmix(st) % ready.size())];
            HIPCHK(hipMemcpyAsync(T(N, T_X) + (size_t)b * P * dn::kInPad,
                                  N.dsX.as<bf16>() + (size_t)slot * P * dn::kInPad, P * dn::kInPad * sizeof(bf16),
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(N.target.as<float>() + (size_t)b * P * 3, N.dsT.as<float>() + (size_t)slot * P * 3,
                                  P * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        dn::launch_input_rgb(T(N, T_X), (long long)batch * (long long)P, N.rgb4.as<bf16>(), s);
        forward(N, s);
        backward(N, s);
        adamStep(N, lr, weightDecay, s);
        dn::launch_copy_bytes(lossHist.as<float>() + it, N.loss.p, sizeof(float), s);
    }
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (losses) HIPCHK(hipMemcpy(losses, lossHist.p, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost));
    return BTE_OK;
}

}  // extern "C"
