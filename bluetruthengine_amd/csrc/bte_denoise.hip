// Denoiser kernels (bf16 MFMA implicit-GEMM convolutions and their
// companions).  Shapes and conventions are documented in bte_denoise.h.
#include "bte_denoise.h"

#include "bte_device.h"

#include <algorithm>
#include <utility>
#include <vector>

namespace bte {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

constexpr int kBlock = 256;

__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ float f(bf16 x) { return (float)x; }

// ---------------------------------------------------------------------------
// 3x3 convolution as an implicit GEMM: M = pixels, N = output channels,
// K = 9 taps x input channels.  A wave owns 32 consecutive pixels (two 16-row
// MFMA tiles) and NT 16-column tiles; each K step takes 32 input channels of
// one tap: lane (row r, group g) loads the 8 channels [8g, 8g + 8) of its
// pixel's tapped neighbour (16 bytes, NHWC) and of its output channel's
// weights, and the MFMA pairs them.  The accumulator lane layout of
// v_mfma_f32_16x16x32_bf16 is D[4g + i][r] in element i.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(kBlock) void k_conv3x3(ConvArgs a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const long long P = (long long)a.n * a.h * a.w_;
    const long long base = ((long long)blockIdx.x * 4 + wave) * 32;
    if (base >= P) return;
    const int coBase = blockIdx.y * NT * 16;
    const int cin = a.c0 + a.c1;
    const int hh = a.h >> 1, wh = a.w_ >> 1;
    int pn[2], py[2], px[2];
    bool pv[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const long long p = base + mt * 16 + r16;
        pv[mt] = p < P;
        const long long pp = pv[mt] ? p : 0;
        px[mt] = (int)(pp % a.w_);
        const long long q = pp / a.w_;
        py[mt] = (int)(q % a.h);
        pn[mt] = (int)(q / a.h);
    }
    f32x4 acc[2][NT];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int t = 0; t < 9; ++t) {
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        const bf16* s0p[2];
        const bf16* s1p[2];
        bool ok[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            ok[mt] = pv[mt] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_;
            const int y0 = ok[mt] ? yy : 0, x0 = ok[mt] ? xx : 0;
            s0p[mt] = a.up0 ? a.src0 + (((size_t)pn[mt] * hh + (y0 >> 1)) * wh + (x0 >> 1)) * a.c0
                            : a.src0 + (((size_t)pn[mt] * a.h + y0) * a.w_ + x0) * a.c0;
            s1p[mt] = a.src1 ? a.src1 + (((size_t)pn[mt] * a.h + y0) * a.w_ + x0) * a.c1 : nullptr;
        }
        const bf16* wrow = a.w + ((size_t)(coBase + r16) * 9 + t) * cin + 8 * g;
        for (int c = 0; c < cin; c += 32) {
            bf16x8 A[2];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const bf16* src = (c < a.c0) ? s0p[mt] + c : s1p[mt] + (c - a.c0);
                A[mt] = ok[mt] ? *reinterpret_cast<const bf16x8*>(src + 8 * g) : zero8();
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const bf16x8 B = *reinterpret_cast<const bf16x8*>(wrow + (size_t)nt * 16 * 9 * cin + c);
                acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B, acc[0][nt], 0, 0, 0);
                acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B, acc[1][nt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long p = base + mt * 16 + 4 * g + i;
            if (p >= P) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = coBase + nt * 16 + r16;
                if (co >= a.cout) continue;
                float v = acc[mt][nt][i];
                if (a.bias) v += a.bias[co];
                if (a.resid) v += f(a.resid[(size_t)p * a.residC + co]);
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.mask && !(f(a.mask[(size_t)p * a.maskC + co]) > 0.0f)) v = 0.0f;
                if (a.dstF) a.dstF[(size_t)p * a.dstC + co] = v;
                else a.dstB[(size_t)p * a.dstC + co] = (bf16)v;
            }
        }
    }
}

// The same convolution with the weight tile of each K step staged in LDS and
// shared by the block's four waves, and four 16-pixel M tiles per wave (64
// pixels), so every weight fragment read from LDS feeds four MFMAs.  The next
// step's weights are fetched into registers while the current step computes.
// LDS rows are padded to 80 bytes to spread the 16-byte reads over the banks.
constexpr int kLdsRow = 40;   // bf16 per staged weight row (32 used)

template <int NT, int MT>
__global__ __launch_bounds__(kBlock) void k_conv3x3_lds(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) bf16 sB[NT * 16 * kLdsRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const long long P = (long long)a.n * a.h * a.w_;
    const long long base = ((long long)blockIdx.x * 4 + wave) * (16 * MT);
    const bool waveActive = base < P;
    const int coBase = blockIdx.y * NT * 16;
    const int cin = a.c0 + a.c1;
    const int hh = a.h >> 1, wh = a.w_ >> 1;
    int pn[MT], py[MT], px[MT];
    bool pv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const long long p = base + mt * 16 + r16;
        pv[mt] = p < P;
        const long long pp = pv[mt] ? p : 0;
        px[mt] = (int)(pp % a.w_);
        const long long q = pp / a.w_;
        py[mt] = (int)(q % a.h);
        pn[mt] = (int)(q / a.h);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    // cooperative staging: 16-byte pieces (row, quarter) of the NT*16 x 32 tile
    constexpr int kPieces = NT * 16 * 4;
    constexpr int kPer = (kPieces + kBlock - 1) / kBlock;
    const int cSteps = cin / 32;
    const int kSteps = 9 * cSteps;
    bf16x8 stage[kPer];
    auto fetch = [&](int ks) {
        const int t = ks / cSteps, c = (ks % cSteps) * 32;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            if (piece < kPieces) {
                const int row = piece >> 2, qt = piece & 3;
                stage[j] = *reinterpret_cast<const bf16x8*>(a.w + ((size_t)(coBase + row) * 9 + t) * cin + c + 8 * qt);
            }
        }
    };
    fetch(0);
    for (int ks = 0; ks < kSteps; ++ks) {
        __syncthreads();                       // every wave is done with the previous tile
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            if (piece < kPieces) {
                const int row = piece >> 2, qt = piece & 3;
                *reinterpret_cast<bf16x8*>(sB + row * kLdsRow + 8 * qt) = stage[j];
            }
        }
        __syncthreads();
        if (ks + 1 < kSteps) fetch(ks + 1);
        if (!waveActive) continue;
        const int t = ks / cSteps, c = (ks % cSteps) * 32;
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        bf16x8 A[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int yy = py[mt] + dy, xx = px[mt] + dx;
            const bool ok = pv[mt] && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_;
            const bf16* src;
            if (c < a.c0) {
                src = a.up0 ? a.src0 + (((size_t)pn[mt] * hh + (yy >> 1)) * wh + (xx >> 1)) * a.c0 + c
                            : a.src0 + (((size_t)pn[mt] * a.h + yy) * a.w_ + xx) * a.c0 + c;
            } else {
                src = a.src1 + (((size_t)pn[mt] * a.h + yy) * a.w_ + xx) * a.c1 + (c - a.c0);
            }
            A[mt] = ok ? *reinterpret_cast<const bf16x8*>(src + 8 * g) : zero8();
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 B = *reinterpret_cast<const bf16x8*>(sB + (nt * 16 + r16) * kLdsRow + 8 * g);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt], B, acc[mt][nt], 0, 0, 0);
        }
    }
    if (!waveActive) return;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long p = base + mt * 16 + 4 * g + i;
            if (p >= P) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = coBase + nt * 16 + r16;
                if (co >= a.cout) continue;
                float v = acc[mt][nt][i];
                if (a.bias) v += a.bias[co];
                if (a.resid) v += f(a.resid[(size_t)p * a.residC + co]);
                if (a.relu) v = fmaxf(v, 0.0f);
                if (a.mask && !(f(a.mask[(size_t)p * a.maskC + co]) > 0.0f)) v = 0.0f;
                if (a.dstF) a.dstF[(size_t)p * a.dstC + co] = v;
                else a.dstB[(size_t)p * a.dstC + co] = (bf16)v;
            }
        }
    }
}

// Epilogue of a tile kernel: bias, residual, ReLU, mask and the store of one
// wave's MT x 16 pixels x NT x 16 channels.  The tap loops take the weights as
// the MFMA's A operand and the pixels as B, so the accumulators are
// transposed: element i of lane (g, r16) in acc[mt][nt] is output channel
// coBase + nt * 16 + 4 g + i of pixel (wave row mt >> 1, column (mt & 1) * 16
// + r16).  A lane therefore stores 4 consecutive channels of one pixel at once
// (8 bytes of bf16) and reads a residual or mask the same way.  EPI 0: plain;
// 1 (EXTRA): a residual or a mask (never both), read beforehand by epiLoad;
// 2: with the 2x2 average pooling of the stored outputs.
__device__ __forceinline__ bf16x4 zero4() { return __builtin_bit_cast(bf16x4, u32x2{0u, 0u}); }

// Residual or mask values of a tile's outputs (4 channels per lane and
// (mt, nt)); issued before the next tile's loads, so that waiting for them
// does not wait for those.
template <int MT, int NT, int RW>
__device__ __forceinline__ void epiLoad(const ConvArgs& a, int n, int x0, int y0, int wave, int g, int r16,
                                        int coBase, bf16x4 (&ext)[MT][NT]) {
    const bf16* src = a.resid ? a.resid : a.mask;
    const int srcC = a.resid ? a.residC : a.maskC;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int y = y0 + wave * RW + (mt >> 1), x = x0 + (mt & 1) * 16 + r16;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = coBase + nt * 16 + 4 * g;
            const bool in = src != nullptr && y < a.h && x < a.w_ && co < a.cout;
            const size_t p = in ? ((size_t)n * a.h + y) * a.w_ + x : 0;
            ext[mt][nt] = in ? *reinterpret_cast<const bf16x4*>(src + p * srcC + co) : zero4();
        }
    }
}

// Biases of the lane's output channels (4 per N tile).
template <int NT>
__device__ __forceinline__ void laneBiases(const ConvArgs& a, int coBase, int g, float (&biasv)[NT][4]) {
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = coBase + nt * 16 + 4 * g + i;
            biasv[nt][i] = (a.bias && co < a.cout) ? a.bias[co] : 0.0f;
        }
}

template <int MT, int NT, int RW, int EPI>
__device__ __forceinline__ void tileEpilogue(const ConvArgs& a, f32x4 (&acc)[MT][NT], const float (&biasv)[NT][4],
                                             const bf16x4 (&ext)[MT][NT], int n, int x0, int y0, int wave, int g,
                                             int r16, int coBase) {
    constexpr bool EXTRA = EPI == 1;
    if constexpr (EPI == 2) {
        // outputs and their 2x2 averages (k_avgpool's sum order over the stored
        // bf16 values): rows 2r, 2r + 1 of the wave are M tiles m0 = 4r + hf and
        // m0 + 2; the horizontal partner of a pixel is the neighbouring lane
        static_assert(RW % 2 == 0, "pooling needs row pairs within a wave");
        const int h2 = a.h >> 1, w2 = a.w_ >> 1;
#pragma unroll
        for (int r = 0; r < RW / 2; ++r) {
            const int y = y0 + wave * RW + 2 * r;
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const int m0 = 4 * r + hf, m1 = m0 + 2;
                const int x = x0 + hf * 16 + r16;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const int co = coBase + nt * 16 + 4 * g;
                    const bool in = y < a.h && x < a.w_ && co < a.cout;   // h, w even: row and column pairs
                    bf16x4 bt, bb;
                    float top[4], bot[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        float v0 = acc[m0][nt][i], v1 = acc[m1][nt][i];
                        if (a.bias) { v0 += biasv[nt][i]; v1 += biasv[nt][i]; }
                        if (a.relu) { v0 = fmaxf(v0, 0.0f); v1 = fmaxf(v1, 0.0f); }
                        bt[i] = (bf16)v0;
                        bb[i] = (bf16)v1;
                        top[i] = f(bt[i]);
                        bot[i] = f(bb[i]);
                    }
                    const size_t p = in ? ((size_t)n * a.h + y) * a.w_ + x : 0;
                    if (in) {
                        *reinterpret_cast<bf16x4*>(a.dstB + p * a.dstC + co) = bt;
                        *reinterpret_cast<bf16x4*>(a.dstB + (p + a.w_) * a.dstC + co) = bb;
                    }
                    float topN[4], botN[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        topN[i] = __shfl_xor(top[i], 1, 64);
                        botN[i] = __shfl_xor(bot[i], 1, 64);
                    }
                    if (in && (r16 & 1) == 0) {
                        bf16x4 pv;
#pragma unroll
                        for (int i = 0; i < 4; ++i) pv[i] = (bf16)(0.25f * (top[i] + topN[i] + bot[i] + botN[i]));
                        *reinterpret_cast<bf16x4*>(a.pool + (((size_t)n * h2 + (y >> 1)) * w2 + (x >> 1)) * a.dstC +
                                                   co) = pv;
                    }
                }
            }
        }
        return;
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int y = y0 + wave * RW + (mt >> 1), x = x0 + (mt & 1) * 16 + r16;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = coBase + nt * 16 + 4 * g;
            if (y >= a.h || x >= a.w_ || co >= a.cout) continue;
            const size_t p = ((size_t)n * a.h + y) * a.w_ + x;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = acc[mt][nt][i];
                if (a.bias) v[i] += biasv[nt][i];
                if (EXTRA && a.resid) v[i] += f(ext[mt][nt][i]);
                if (a.relu) v[i] = fmaxf(v[i], 0.0f);
                if (EXTRA && a.mask && !(f(ext[mt][nt][i]) > 0.0f)) v[i] = 0.0f;
            }
            if (a.dstF) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (co + i < a.cout) a.dstF[p * a.dstC + co + i] = v[i];
            } else {
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; ++i) o[i] = (bf16)v[i];
                *reinterpret_cast<bf16x4*>(a.dstB + p * a.dstC + co) = o;
            }
        }
    }
}

// The convolution as a spatial tile per block: 4 * RW image rows x 32 pixels
// (each wave RW rows, i.e. 2 * RW M tiles of 16 pixels) x NT * 16 output
// channels.  For every 32-channel slice of the input, the block stages in LDS
// the tile with its one-pixel halo (zero outside the image; nearest
// upsampling and the channel concatenation resolved while staging) and the
// slice's weights for all 9 taps; the 9 taps then read their A fragments
// from shifted windows of that one tile, so activations cross the memory
// hierarchy once per slice instead of once per tap.  LDS rows (one pixel's or
// one output channel's 32 channels) are padded to 80 bytes to spread the
// 16-byte fragment reads over the banks.
constexpr int kTileW = 32;   // output pixels per tile row

// LDS layout of a staged input tile or weight slice: one row (a pixel's or
// an output channel's 32 bf16 of the slice) is 64 bytes, and its 16-byte
// chunk q sits at chunk q ^ ((row >> 1) & 3).  Measured on this LDS, both
// access patterns of the tile kernels then run without bank conflicts: the
// staging stores (4 consecutive lanes write the 4 chunks of one row, a wave
// 16 consecutive rows) and the MFMA fragment reads (16 consecutive rows from
// any first row, one chunk per 16 lanes), where rows of 80 bytes (32 + 8
// bf16) took twice the cycles per fragment read and 4/3 per store.
constexpr int kTileRow = 32;   // bf16 per staged row
__device__ __forceinline__ int ldsAt(int row, int q) { return row * kTileRow + 8 * (q ^ ((row >> 1) & 3)); }
// Staging order: piece p is chunk p & 3 of row p >> 2.
constexpr int stagedPieces(int rows) { return rows * 4; }
__device__ __forceinline__ int pieceRow(int piece) { return piece >> 2; }
__device__ __forceinline__ int pieceQ(int piece) { return piece & 3; }

// Per-lane element offsets of the MFMA fragment reads from a staged TR x TC
// tile.  Lane (g, r16) of a wave reads chunk g of row
// (wave * RW + m + oy) * TC + h * 16 + r16 + ox (m: M-tile row, h: M-tile
// half, ox / oy: tap offset + 1).  With TC and kTileW / 2 even multiples
// as asserted, the row's swizzle key is ((wave * RW + m + oy) * TC / 2 +
// ((r16 + ox) >> 1)) & 3, so a read is
//   (((m + oy) * TC + h * 16) * kTileRow) + aOff[ox][(m + oy) & 3]
// with the lane part precomputed here (12 registers, no address arithmetic
// in the tap loops).  Weight rows of a slice start at multiples of 16, so
// their key is (r16 >> 1) & 3 and bOff = ldsAt(r16, g).
template <int RW, int TC>
__device__ __forceinline__ void fragOffsets(int wave, int g, int r16, int (&aOff)[3][4], int& bOff) {
    static_assert(TC % 2 == 0 && (kTileW / 2) % 4 == 0, "swizzle key of a fragment row");
#pragma unroll
    for (int ox = 0; ox < 3; ++ox)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int key = (((wave * RW + c) * (TC / 2)) + ((r16 + ox) >> 1)) & 3;
            aOff[ox][c] = (wave * RW * TC + r16 + ox) * kTileRow + 8 * (g ^ key);
        }
    bOff = ldsAt(r16, g);
}
template <int NT, int RW>
__global__ __launch_bounds__(kBlock) void k_conv3x3_tile(ConvArgs a, int tilesX, int tilesY) {
    constexpr int TH = 4 * RW, MT = 2 * RW;
    constexpr int TR = TH + 2, TC = kTileW + 2;
    constexpr int kRowsA = TR * TC, kPiecesA = stagedPieces(kRowsA), kPerA = (kPiecesA + kBlock - 1) / kBlock;
    constexpr int kPiecesW = stagedPieces(9 * NT * 16), kPerW = (kPiecesW + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) bf16 sA[TR * TC * kTileRow];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * NT * 16 * kTileRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    int aOff[3][4], bOff;
    fragOffsets<RW, TC>(wave, g, r16, aOff, bOff);
    int bid = blockIdx.x;
    const int tx = bid % tilesX;
    bid /= tilesX;
    const int ty = bid % tilesY, n = bid / tilesY;
    const int x0 = tx * kTileW, y0 = ty * TH;
    const int coBase = blockIdx.y * NT * 16;
    const int cin = a.c0 + a.c1;
    const int hh = a.h >> 1, wh = a.w_ >> 1;
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int c = 0; c < cin; c += 32) {
        if (c) __syncthreads();                // every wave is done with the previous slice
        {   // input tile: global -> registers -> LDS
            bf16x8 va[kPerA];
#pragma unroll
            for (int j = 0; j < kPerA; ++j) {
                const int piece = threadIdx.x + j * kBlock;
                va[j] = zero8();
                const int sidx = pieceRow(piece), q = pieceQ(piece);
                if (piece < kPiecesA && sidx < kRowsA) {
                    const int yy = y0 - 1 + sidx / TC, xx = x0 - 1 + sidx % TC;
                    if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_) {
                        const bf16* src;
                        if (c < a.c0)
                            src = a.up0 ? a.src0 + (((size_t)n * hh + (yy >> 1)) * wh + (xx >> 1)) * a.c0 + c
                                        : a.src0 + (((size_t)n * a.h + yy) * a.w_ + xx) * a.c0 + c;
                        else
                            src = a.src1 + (((size_t)n * a.h + yy) * a.w_ + xx) * a.c1 + (c - a.c0);
                        va[j] = *reinterpret_cast<const bf16x8*>(src + 8 * q);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kPerA; ++j) {
                const int piece = threadIdx.x + j * kBlock;
                const int row = pieceRow(piece);
                if (piece < kPiecesA && row < kRowsA)
                    *reinterpret_cast<bf16x8*>(sA + ldsAt(row, pieceQ(piece))) = va[j];
            }
        }
        {   // weight slice of all 9 taps: global -> registers -> LDS
            bf16x8 vw[kPerW];
#pragma unroll
            for (int j = 0; j < kPerW; ++j) {
                const int piece = threadIdx.x + j * kBlock;
                if (piece < kPiecesW) {
                    const int row = pieceRow(piece), q = pieceQ(piece);
                    const int t = row / (NT * 16), r = row % (NT * 16);
                    vw[j] = *reinterpret_cast<const bf16x8*>(a.w + ((size_t)(coBase + r) * 9 + t) * cin + c + 8 * q);
                }
            }
#pragma unroll
            for (int j = 0; j < kPerW; ++j) {
                const int piece = threadIdx.x + j * kBlock;
                if (piece < kPiecesW) *reinterpret_cast<bf16x8*>(sW + ldsAt(pieceRow(piece), pieceQ(piece))) = vw[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int oy = t / 3, ox = t % 3;   // tap offset + 1: tile row / column shift
            bf16x8 B[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                B[nt] = *reinterpret_cast<const bf16x8*>(sW + (t * NT + nt) * 16 * kTileRow + bOff);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int m = (mt >> 1) + oy;
                const bf16x8 A = *reinterpret_cast<const bf16x8*>(sA + (m * TC + (mt & 1) * 16) * kTileRow + aOff[ox][m & 3]);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[nt], A, acc[mt][nt], 0, 0, 0);
            }
        }
    }
    float biasv[NT][4];
    laneBiases<NT>(a, coBase, g, biasv);
    bf16x4 ext[MT][NT];
    if (a.resid || a.mask) {
        epiLoad<MT, NT, RW>(a, n, x0, y0, wave, g, r16, coBase, ext);
        tileEpilogue<MT, NT, RW, 1>(a, acc, biasv, ext, n, x0, y0, wave, g, r16, coBase);
    } else {
        tileEpilogue<MT, NT, RW, 0>(a, acc, biasv, ext, n, x0, y0, wave, g, r16, coBase);
    }
}

// Loads of one 32-channel slice of a tile with its halo, without branches:
// piece p of the tile (pixel pieceRow(p) of the TR x TC window, 8 channels
// pieceQ(p)) reads from a clamped, in-image address and keeps zero outside the image, so
// every lane issues its loads back to back.  The slice lies in src0 (at half
// resolution when up0) or in src1.
struct SliceSrc {
    const bf16* base;   // first channel of the slice
    int stride;         // channels per pixel of that tensor
    int sh;             // 1: read through the nearest 2x upsample
};
__device__ __forceinline__ SliceSrc sliceSrc(const ConvArgs& a, int c) {
    SliceSrc r;
    const bool s0 = c < a.c0;
    r.base = s0 ? a.src0 + c : a.src1 + (c - a.c0);
    r.stride = s0 ? a.c0 : a.c1;
    r.sh = (s0 && a.up0) ? 1 : 0;
    return r;
}
template <int TC, int kRows, int kPer>
__device__ __forceinline__ void loadSlice(const ConvArgs& a, const SliceSrc& src, int n, int x0, int y0,
                                          bf16x8 (&va)[kPer]) {
    const int hS = a.h >> src.sh, wS = a.w_ >> src.sh;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int piece = threadIdx.x + j * kBlock;
        const int row = pieceRow(piece), q = pieceQ(piece);
        const bool valid = piece < stagedPieces(kRows) && row < kRows;
        const int sidx = valid ? row : kRows - 1;
        const int yy = y0 - 1 + sidx / TC, xx = x0 - 1 + sidx % TC;
        const bool in = valid && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_;
        const int yc = min(max(yy, 0), a.h - 1), xc = min(max(xx, 0), a.w_ - 1);
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(
            src.base + (((size_t)n * hS + (yc >> src.sh)) * wS + (xc >> src.sh)) * src.stride + 8 * q);
        va[j] = in ? v : zero8();
    }
}

// The 9 taps of one staged slice: per tap the weight fragments of the lane's
// output channels, then per M tile its pixel fragment and NT MFMAs (weights
// as the A operand: transposed accumulators, see tileEpilogue).
template <int MT, int NT, int RW, int TC>
__device__ __forceinline__ void tapsFromLds(const bf16* sA, const bf16* sW, const int (&aOff)[3][4], int bOff,
                                            f32x4 (&acc)[MT][NT]) {
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        const int oy = t / 3, ox = t % 3;   // tap offset + 1: tile row / column shift
        bf16x8 B[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            B[nt] = *reinterpret_cast<const bf16x8*>(sW + (t * NT + nt) * 16 * kTileRow + bOff);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int m = (mt >> 1) + oy;
            const bf16x8 A = *reinterpret_cast<const bf16x8*>(sA + (m * TC + (mt & 1) * 16) * kTileRow + aOff[ox][m & 3]);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(B[nt], A, acc[mt][nt], 0, 0, 0);
        }
    }
}

// Single-slice layers (32 input channels): persistent blocks keep the weights
// of all 9 taps in LDS for their whole run and walk tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...; the next tile's input (with halo) is loaded
// into registers while the current one is computed and stored.  Registers
// are held to 3 waves per SIMD (168), which the 32-channel layers reach
// without spilling (measured faster than 2 waves per SIMD at 172).
template <int NT, int RW, int EPI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_conv3x3_tile1(ConvArgs a, int tilesX,
                                                                                             int tilesY, int numTiles) {
    constexpr int TH = 4 * RW, MT = 2 * RW;
    constexpr int TR = TH + 2, TC = kTileW + 2;
    constexpr int kRowsA = TR * TC, kPiecesA = stagedPieces(kRowsA), kPerA = (kPiecesA + kBlock - 1) / kBlock;
    constexpr int kPiecesW = stagedPieces(9 * NT * 16), kPerW = (kPiecesW + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) bf16 sA[TR * TC * kTileRow];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * NT * 16 * kTileRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    int aOff[3][4], bOff;
    fragOffsets<RW, TC>(wave, g, r16, aOff, bOff);
    const int coBase = blockIdx.y * NT * 16;
    {   // weights of the one 32-channel slice, all taps
        bf16x8 vw[kPerW];
#pragma unroll
        for (int j = 0; j < kPerW; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            if (piece < kPiecesW) {
                const int row = pieceRow(piece), q = pieceQ(piece);
                const int t = row / (NT * 16), r = row % (NT * 16);
                vw[j] = *reinterpret_cast<const bf16x8*>(a.w + ((size_t)(coBase + r) * 9 + t) * 32 + 8 * q);
            }
        }
#pragma unroll
        for (int j = 0; j < kPerW; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            if (piece < kPiecesW) *reinterpret_cast<bf16x8*>(sW + ldsAt(pieceRow(piece), pieceQ(piece))) = vw[j];
        }
    }
    // per-lane output channels' biases, read once (no global loads in the
    // epilogue of bias-only layers, so the next tile's loads stay in flight)
    float biasv[NT][4];
    laneBiases<NT>(a, coBase, g, biasv);
    bf16x8 va[kPerA];
    const SliceSrc src = sliceSrc(a, 0);
    auto loadTile = [&](int tile) {
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int ty = rest % tilesY, n = rest / tilesY;
        loadSlice<TC, kRowsA, kPerA>(a, src, n, tx * kTileW, ty * TH, va);
    };
    int tile = blockIdx.x;
    if (tile < numTiles) loadTile(tile);
    for (; tile < numTiles; tile += gridDim.x) {
        __syncthreads();                       // the previous tile's taps are done with sA
#pragma unroll
        for (int j = 0; j < kPerA; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            const int row = pieceRow(piece);
            if (piece < kPiecesA && row < kRowsA) *reinterpret_cast<bf16x8*>(sA + ldsAt(row, pieceQ(piece))) = va[j];
        }
        __syncthreads();
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int ty = rest % tilesY, n = rest / tilesY;
        bf16x4 ext[MT][NT];
        if constexpr (EPI == 1) epiLoad<MT, NT, RW>(a, n, tx * kTileW, ty * TH, wave, g, r16, coBase, ext);
        const int next = tile + (int)gridDim.x;
        if (next < numTiles) loadTile(next);   // in flight while this tile is computed
        f32x4 acc[MT][NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        tapsFromLds<MT, NT, RW, TC>(sA, sW, aOff, bOff, acc);
        tileEpilogue<MT, NT, RW, EPI>(a, acc, biasv, ext, n, tx * kTileW, ty * TH, wave, g, r16, coBase);
    }
}

// Multi-slice layers (64 or more input channels): persistent blocks walk
// (tile, 32-channel slice) items in order; while one slice's 9 taps run from
// LDS, the next item's input slice (with halo) and weight slice are loaded
// into registers.  The output of a tile is written after its last slice.
template <int NT, int RW, int EPI>
__global__ __launch_bounds__(kBlock) void k_conv3x3_tileN(ConvArgs a, int tilesX, int tilesY, int numTiles) {
    constexpr int TH = 4 * RW, MT = 2 * RW;
    constexpr int TR = TH + 2, TC = kTileW + 2;
    constexpr int kRowsA = TR * TC, kPiecesA = stagedPieces(kRowsA), kPerA = (kPiecesA + kBlock - 1) / kBlock;
    constexpr int kPiecesW = stagedPieces(9 * NT * 16), kPerW = (kPiecesW + kBlock - 1) / kBlock;
    __shared__ __attribute__((aligned(16))) bf16 sA[TR * TC * kTileRow];
    __shared__ __attribute__((aligned(16))) bf16 sW[9 * NT * 16 * kTileRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    int aOff[3][4], bOff;
    fragOffsets<RW, TC>(wave, g, r16, aOff, bOff);
    const int coBase = blockIdx.y * NT * 16;
    const int cin = a.c0 + a.c1, slices = cin / 32;
    float biasv[NT][4];
    laneBiases<NT>(a, coBase, g, biasv);
    bf16x8 va[kPerA], vw[kPerW];
    auto load = [&](int tile, int slice) {
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int ty = rest % tilesY, n = rest / tilesY;
        const int c = slice * 32;
        loadSlice<TC, kRowsA, kPerA>(a, sliceSrc(a, c), n, tx * kTileW, ty * TH, va);
#pragma unroll
        for (int j = 0; j < kPerW; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            const int pc = piece < kPiecesW ? piece : kPiecesW - 1;
            const int row = pieceRow(pc), q = pieceQ(pc);
            const int t = row / (NT * 16), r = row % (NT * 16);
            vw[j] = *reinterpret_cast<const bf16x8*>(a.w + ((size_t)(coBase + r) * 9 + t) * cin + c + 8 * q);
        }
    };
    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    int tile = blockIdx.x, slice = 0;
    if (tile < numTiles) load(tile, 0);
    while (tile < numTiles) {
        __syncthreads();                       // the previous slice's taps are done with LDS
#pragma unroll
        for (int j = 0; j < kPerA; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            const int row = pieceRow(piece);
            if (piece < kPiecesA && row < kRowsA) *reinterpret_cast<bf16x8*>(sA + ldsAt(row, pieceQ(piece))) = va[j];
        }
#pragma unroll
        for (int j = 0; j < kPerW; ++j) {
            const int piece = threadIdx.x + j * kBlock;
            if (piece < kPiecesW) *reinterpret_cast<bf16x8*>(sW + ldsAt(pieceRow(piece), pieceQ(piece))) = vw[j];
        }
        __syncthreads();
        const bool last = slice + 1 == slices;
        const int nTile = last ? tile + (int)gridDim.x : tile, nSlice = last ? 0 : slice + 1;
        const int tx = tile % tilesX, rest = tile / tilesX;
        const int ty = rest % tilesY, n = rest / tilesY;
        bf16x4 ext[MT][NT];
        if constexpr (EPI == 1)
            if (last) epiLoad<MT, NT, RW>(a, n, tx * kTileW, ty * TH, wave, g, r16, coBase, ext);
        if (nTile < numTiles) load(nTile, nSlice);   // in flight while this slice is computed
        tapsFromLds<MT, NT, RW, TC>(sA, sW, aOff, bOff, acc);
        if (last) {
            tileEpilogue<MT, NT, RW, EPI>(a, acc, biasv, ext, n, tx * kTileW, ty * TH, wave, g, r16, coBase);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
        tile = nTile;
        slice = nSlice;
    }
}

// Blocks of `kernel` (256 threads) resident on the whole device at once, from
// the occupancy calculator; cached per kernel.
static int residentBlocks(const void* kernel) {
    static std::vector<std::pair<const void*, int>> cache;
    for (const auto& kv : cache)
        if (kv.first == kernel) return kv.second;
    int perCU = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kBlock, 0) != hipSuccess || perCU < 1 || cus < 1) {
        perCU = 2;
        cus = 256;
    }
    cache.emplace_back(kernel, perCU * cus);
    return perCU * cus;
}

static int g_convVariant = 2;   // 0: direct loads, 1: LDS-staged weights, 2: LDS-staged tiles
void set_conv_variant(int v) { g_convVariant = v; }
int conv_variant() { return g_convVariant; }

// Launches the convolution; returns whether its kernel also wrote a.pool.
static bool convKernels(const ConvArgs& a, hipStream_t s) {
    const long long P = (long long)a.n * a.h * a.w_;
    const int tiles = a.coutPad / 16;
    if (g_convVariant == 2) {
        // 8 x 32-pixel tiles; at most 64 output channels per block (LDS)
        const int nt4 = (tiles % 4 == 0) ? 4 : (tiles % 2 == 0) ? 2 : 1;
        const int tilesX = (a.w_ + kTileW - 1) / kTileW, tilesY = (a.h + 7) / 8;
        const long long numTiles = (long long)tilesX * tilesY * a.n;
        // persistent blocks: as many as are resident at once, split over the
        // output-channel groups
        auto persistent = [&](const void* kernel, int groups) {
            const long long per = std::max(1, residentBlocks(kernel) / groups);
            return dim3((unsigned)std::max<long long>(1, std::min<long long>(numTiles, per)), (unsigned)groups);
        };
        if (a.c0 + a.c1 == 32) {
            const bool extra = a.resid || a.mask;
            const bool pool = !extra && a.pool && a.dstB;
#define BTE_TILE1(NT_, EPI_)                                                                                   \
    hipLaunchKernelGGL((k_conv3x3_tile1<NT_, 2, EPI_>),                                                         \
                       persistent(reinterpret_cast<const void*>(&k_conv3x3_tile1<NT_, 2, EPI_>), tiles / NT_), \
                       dim3(kBlock), 0, s, a, tilesX, tilesY, (int)numTiles)
            switch (nt4) {
                case 4: if (extra) BTE_TILE1(4, 1); else BTE_TILE1(4, 0); break;
                case 2:
                    if (extra) BTE_TILE1(2, 1);
                    else if (pool) BTE_TILE1(2, 2);
                    else BTE_TILE1(2, 0);
                    return pool;
                default: if (extra) BTE_TILE1(1, 1); else BTE_TILE1(1, 0); break;
            }
#undef BTE_TILE1
            return false;
        }
        if ((a.c0 + a.c1) >= 192 && nt4 == 4) {
            // many slices and >= 64 output channels: one pass over the input
            // per 64 channels (measured faster than the prefetching kernel's
            // 32-channel groups and than 4-row tiles with 64 channels)
            const dim3 grid((unsigned)numTiles, (unsigned)(tiles / 4));
            hipLaunchKernelGGL((k_conv3x3_tile<4, 2>), grid, dim3(kBlock), 0, s, a, tilesX, tilesY);
            return false;
        }
        // several slices: at most 32 output channels per block (registers for the prefetch)
        const int nt2 = (tiles % 2 == 0) ? 2 : 1;
        const bool extra = a.resid || a.mask;
        const bool pool = !extra && a.pool && a.dstB;
#define BTE_TILEN(NT_, EPI_)                                                                                   \
    hipLaunchKernelGGL((k_conv3x3_tileN<NT_, 2, EPI_>),                                                         \
                       persistent(reinterpret_cast<const void*>(&k_conv3x3_tileN<NT_, 2, EPI_>), tiles / NT_), \
                       dim3(kBlock), 0, s, a, tilesX, tilesY, (int)numTiles)
        if (nt2 == 2) {
            if (extra) BTE_TILEN(2, 1);
            else if (pool) BTE_TILEN(2, 2);
            else BTE_TILEN(2, 0);
            return pool;
        }
        if (extra) BTE_TILEN(1, 1); else BTE_TILEN(1, 0);
#undef BTE_TILEN
        return false;
    }
    const int nt = (tiles % 8 == 0) ? 8 : (tiles % 4 == 0) ? 4 : (tiles % 2 == 0) ? 2 : 1;
    if (g_convVariant == 1) {
        // pixels per block: 4 waves x MT x 16 (MT = 2 for 128-wide N tiles)
        const unsigned gy = (unsigned)(tiles / nt);
        switch (nt) {
            case 8: hipLaunchKernelGGL((k_conv3x3_lds<8, 2>), dim3((unsigned)((P + 127) / 128), gy), dim3(kBlock), 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_conv3x3_lds<4, 4>), dim3((unsigned)((P + 255) / 256), gy), dim3(kBlock), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_conv3x3_lds<2, 4>), dim3((unsigned)((P + 255) / 256), gy), dim3(kBlock), 0, s, a); break;
            default: hipLaunchKernelGGL((k_conv3x3_lds<1, 4>), dim3((unsigned)((P + 255) / 256), gy), dim3(kBlock), 0, s, a); break;
        }
        return false;
    }
    const dim3 grid((unsigned)((P + 127) / 128), (unsigned)(tiles / nt));
    switch (nt) {
        case 8: hipLaunchKernelGGL(k_conv3x3<8>, grid, dim3(kBlock), 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_conv3x3<4>, grid, dim3(kBlock), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_conv3x3<2>, grid, dim3(kBlock), 0, s, a); break;
        default: hipLaunchKernelGGL(k_conv3x3<1>, grid, dim3(kBlock), 0, s, a); break;
    }
    return false;
}

void launch_conv3x3(const ConvArgs& a, hipStream_t s) {
    if (!convKernels(a, s) && a.pool && a.dstB) launch_avgpool(a.dstB, a.pool, a.n, a.h, a.w_, a.dstC, s);
}

// ---------------------------------------------------------------------------
// Weight gradient as a GEMM over pixels: dW_tap (cout x cin) = G^T X_tap with
// K = pixels.  NHWC tensors keep a pixel's channels together, while an MFMA
// operand wants 8 consecutive pixels of one channel per lane, so a block
// stages each K step (32 pixels of one image row) transposed in LDS: the
// gradient rows G^T (CO channels x 32 pixels) and, for each of the 9 taps, the
// shifted input window X_tap (32 channels x 32 pixels, zero outside the image;
// nearest upsampling and the channel concatenation resolved while staging).
// The input rows y-1..y+1 are read once per step and written to the three
// horizontal shifts.  Each wave then runs its (tap, m tile, n tile) 16 x 16
// tiles from LDS, the A fragment of an m tile shared by its taps.  A block
// covers CO = 16 MT output channels x 32 input channels x all 9 taps over a
// chunk of K steps; its partial sums go to part[chunk] and a second kernel
// adds the chunks in order (deterministic, no atomics).
// ---------------------------------------------------------------------------
constexpr int kWgRow = 40;   // bf16 per staged LDS row (32 pixels used): 80-byte rows spread the 16-byte reads
template <int MT>
__global__ __launch_bounds__(kBlock) void k_wgrad3x3(WgradArgs a, int stepsPerChunk, long long steps) {
    constexpr int CO = 16 * MT;
    constexpr int TILES = 9 * MT * 2;
    constexpr int PER_WAVE = (TILES + 3) / 4;
    __shared__ __attribute__((aligned(16))) bf16 sG[CO * kWgRow];
    __shared__ __attribute__((aligned(16))) bf16 sX[9 * 32 * kWgRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, r16 = lane & 15;
    const int cin = a.c0 + a.c1;
    const int coBase = blockIdx.y * CO;
    const int ciBase = blockIdx.z * 32;
    const bool fromS0 = ciBase < a.c0;
    const int hh = a.h >> 1, wh = a.w_ >> 1;
    const int rowSteps = a.w_ / 32;
    const long long st0 = (long long)blockIdx.x * stepsPerChunk;
    const long long st1 = st0 + stepsPerChunk < steps ? st0 + stepsPerChunk : steps;
    f32x4 acc[PER_WAVE];
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // bias gradient (blocks of the first input-channel slice): thread v < 4 CO
    // stages channels [8 (v % (CO / 8)), + 8) of pixel v / (CO / 8) every step
    // and sums them
    const bool biasSums = a.dB != nullptr && blockIdx.z == 0;
    float bsum[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (long long st = st0; st < st1; ++st) {
        const int x0 = (int)(st % rowSteps) * 32;
        const long long q = st / rowSteps;
        const int y = (int)(q % a.h), n = (int)(q / a.h);
        const long long pRow = ((long long)n * a.h + y) * a.w_ + x0;
        // G^T: CO channels x 32 pixels
        for (int v = threadIdx.x; v < 32 * (CO / 8); v += kBlock) {
            const int k = v / (CO / 8), c8 = v % (CO / 8);
            const int co = coBase + 8 * c8;
            bf16x8 x = zero8();
            if (co < a.gC) x = *reinterpret_cast<const bf16x8*>(a.g + (size_t)(pRow + k) * a.gC + co);
#pragma unroll
            for (int e = 0; e < 8; ++e) sG[(8 * c8 + e) * kWgRow + k] = x[e];
            if (biasSums) {   // 32 * CO / 8 <= kBlock: v == threadIdx.x, the same (k, c8) every step
#pragma unroll
                for (int e = 0; e < 8; ++e) bsum[e] += (float)x[e];
            }
        }
        // X: rows y-1..y+1, columns x0-1..x0+32, 32 channels, into the 9 tap windows
        for (int v = threadIdx.x; v < 3 * 34 * 4; v += kBlock) {
            const int c8 = v & 3, j = (v >> 2) % 34, dyi = (v >> 2) / 34;
            const int yy = y + dyi - 1, xx = x0 + j - 1;
            bf16x8 x = zero8();
            if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w_) {
                if (fromS0) {
                    const int c = ciBase + 8 * c8;
                    x = a.up0 ? *reinterpret_cast<const bf16x8*>(
                                    a.src0 + (((size_t)n * hh + (yy >> 1)) * wh + (xx >> 1)) * a.c0 + c)
                              : *reinterpret_cast<const bf16x8*>(
                                    a.src0 + (((size_t)n * a.h + yy) * a.w_ + xx) * a.c0 + c);
                } else {
                    const int c = ciBase - a.c0 + 8 * c8;
                    x = *reinterpret_cast<const bf16x8*>(a.src1 + (((size_t)n * a.h + yy) * a.w_ + xx) * a.c1 + c);
                }
            }
            // tap (dyi, dxi) at pixel k reads column k + dxi
#pragma unroll
            for (int dxi = 0; dxi < 3; ++dxi) {
                const int k = j - dxi;
                if (k < 0 || k >= 32) continue;
                bf16* row = sX + ((dyi * 3 + dxi) * 32 + 8 * c8) * kWgRow + k;
#pragma unroll
                for (int e = 0; e < 8; ++e) row[e * kWgRow] = x[e];
            }
        }
        __syncthreads();
        const bf16x8 A0 = *reinterpret_cast<const bf16x8*>(sG + r16 * kWgRow + 8 * g);
        const bf16x8 A1 = *reinterpret_cast<const bf16x8*>(sG + ((MT - 1) * 16 + r16) * kWgRow + 8 * g);
#pragma unroll
        for (int i = 0; i < PER_WAVE; ++i) {
            const int j = wave + 4 * i;
            if (j < TILES) {
                const int tap = j / (2 * MT), mt = (j >> 1) % MT, nt = j & 1;
                const bf16x8 B = *reinterpret_cast<const bf16x8*>(sX + (tap * 32 + nt * 16 + r16) * kWgRow + 8 * g);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mt == 0 ? A0 : A1, B, acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[4g + e][r16] of tile (tap, mt, nt): output channel coBase + 16 mt + 4g + e, input channel ciBase + 16 nt + r16
    const size_t nw = (size_t)a.coutPad * 9 * cin;
    float* part = a.part + (size_t)blockIdx.x * (nw + a.coutPad);
    if (biasSums) {
        // fold the 32 pixels of each channel group in a fixed order; sX is free now
        float* red = reinterpret_cast<float*>(sX);
        constexpr int V = 32 * (CO / 8);
        if (threadIdx.x < V)
#pragma unroll
            for (int e = 0; e < 8; ++e) red[e * V + threadIdx.x] = bsum[e];
        __syncthreads();
        if (threadIdx.x < CO) {
            const int c8 = threadIdx.x / 8, e = threadIdx.x % 8;
            float t = 0.0f;
            for (int k = 0; k < 32; ++k) t += red[e * V + k * (CO / 8) + c8];
            part[nw + coBase + threadIdx.x] = t;
        }
    }
#pragma unroll
    for (int i = 0; i < PER_WAVE; ++i) {
        const int j = wave + 4 * i;
        if (j < TILES) {
            const int tap = j / (2 * MT), mt = (j >> 1) % MT, nt = j & 1;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int co = coBase + mt * 16 + 4 * g + e;
                part[((size_t)co * 9 + tap) * cin + ciBase + nt * 16 + r16] = acc[i][e];
            }
        }
    }
}

// Chunk sums in two passes with a fixed order: pass 1 adds the chunks of group
// y (chunks [y cpg, (y + 1) cpg)) of element i into part2[y][i]; pass 2 adds
// the groups into the weight (i < nW: element i of dW, rows below cout) or
// bias (i - nW) gradient.  Element i lies at offset i (weights) or coutPad 9
// cin + i - nW (biases) of a chunk's partial sums.
__global__ __launch_bounds__(kBlock) void k_wgrad_reduce1(const float* part, int chunks, int cpg, long long stride,
                                                          long long nW, long long biasOff, long long count,
                                                          float* part2) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const long long off = i < nW ? i : biasOff + (i - nW);
    const int c0 = blockIdx.y * cpg, c1 = min(c0 + cpg, chunks);
    float s = 0.0f;
    for (int c = c0; c < c1; ++c) s += part[(size_t)c * stride + off];
    part2[(size_t)blockIdx.y * count + i] = s;
}
__global__ __launch_bounds__(kBlock) void k_wgrad_reduce2(const float* part2, int groups, long long nW,
                                                          long long count, float* dW, float* dB) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int y = 0; y < groups; ++y) s += part2[(size_t)y * count + i];
    if (i < nW) dW[i] += s;
    else dB[i - nW] += s;
}

// K steps (32-pixel row segments) per chunk: enough blocks to fill the device
// about twice, and at most kWgradPartFloats of partial sums.
static void wgradChunks(const WgradArgs& a, int& chunks, int& stepsPerChunk) {
    const long long steps = (long long)a.n * a.h * (a.w_ / 32);
    const int mt = a.coutPad % 32 == 0 ? 2 : 1;
    const long long tiles = (long long)(a.coutPad / (16 * mt)) * ((a.c0 + a.c1) / 32);
    const long long perChunk = (long long)a.coutPad * 9 * (a.c0 + a.c1) + a.coutPad;
    long long c = std::max(1ll, 2048 / tiles);
    c = std::min(c, std::max(1ll, (long long)kWgradPartFloats / perChunk));
    c = std::min(c, steps);
    stepsPerChunk = (int)((steps + c - 1) / c);
    chunks = (int)((steps + stepsPerChunk - 1) / stepsPerChunk);
}

void launch_wgrad3x3(const WgradArgs& a, hipStream_t s) {
    const long long steps = (long long)a.n * a.h * (a.w_ / 32);
    const int cin = a.c0 + a.c1;
    int chunks = 0, spc = 0;
    wgradChunks(a, chunks, spc);
    if (a.coutPad % 32 == 0)
        hipLaunchKernelGGL(k_wgrad3x3<2>, dim3(chunks, a.coutPad / 32, cin / 32), dim3(kBlock), 0, s, a, spc, steps);
    else
        hipLaunchKernelGGL(k_wgrad3x3<1>, dim3(chunks, a.coutPad / 16, cin / 32), dim3(kBlock), 0, s, a, spc, steps);
    const long long nW = (long long)a.cout * 9 * cin;
    const long long count = nW + (a.dB ? a.cout : 0);
    const long long biasOff = (long long)a.coutPad * 9 * cin;
    const int groups = std::min(chunks, kWgradGroups);
    const int cpg = (chunks + groups - 1) / groups;
    const int used = (chunks + cpg - 1) / cpg;
    float* part2 = a.part + kWgradPartFloats;
    const unsigned xb = (unsigned)((count + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_wgrad_reduce1, dim3(xb, used), dim3(kBlock), 0, s, a.part, chunks, cpg, biasOff + a.coutPad,
                       nW, biasOff, count, part2);
    hipLaunchKernelGGL(k_wgrad_reduce2, dim3(xb), dim3(kBlock), 0, s, part2, used, nW, count, a.dW, a.dB);
}

// ---------------------------------------------------------------------------
// pooling, upsampling gradients
// ---------------------------------------------------------------------------
// One thread per output pixel and 8 channels (16-byte loads and store); each
// channel is (a + b + c + d) * 0.25 in that order.
__global__ void k_avgpool(const bf16* src, bf16* dst, int n, int h, int w, int c) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h2 = h >> 1, w2 = w >> 1, c8 = c >> 3;
    const long long total = (long long)n * h2 * w2 * c8;
    if (i >= total) return;
    const int ch = (int)(i % c8) * 8;
    const long long q = i / c8;
    const int x = (int)(q % w2);
    const long long r = q / w2;
    const int y = (int)(r % h2), nn = (int)(r / h2);
    const size_t b00 = (((size_t)nn * h + 2 * y) * w + 2 * x) * c + ch;
    const bf16x8 v00 = *reinterpret_cast<const bf16x8*>(src + b00);
    const bf16x8 v01 = *reinterpret_cast<const bf16x8*>(src + b00 + c);
    const bf16x8 v10 = *reinterpret_cast<const bf16x8*>(src + b00 + (size_t)w * c);
    const bf16x8 v11 = *reinterpret_cast<const bf16x8*>(src + b00 + (size_t)w * c + c);
    bf16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) o[e] = (bf16)(0.25f * (f(v00[e]) + f(v01[e]) + f(v10[e]) + f(v11[e])));
    *reinterpret_cast<bf16x8*>(dst + q * c + ch) = o;
}

void launch_avgpool(const bf16* src, bf16* dst, int n, int h, int w, int c, hipStream_t s) {
    const long long total = (long long)n * (h / 2) * (w / 2) * (c / 8);
    hipLaunchKernelGGL(k_avgpool, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, n, h, w, c);
}

__global__ void k_upsample_bwd(const bf16* g, int gC, const bf16* act, bf16* out, int n, int h, int w, int ca) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int h2 = h >> 1, w2 = w >> 1;
    const long long total = (long long)n * h2 * w2 * ca;
    if (i >= total) return;
    const int ch = (int)(i % ca);
    const long long q = i / ca;
    const int x = (int)(q % w2);
    const long long r = q / w2;
    const int y = (int)(r % h2), nn = (int)(r / h2);
    const size_t b00 = (((size_t)nn * h + 2 * y) * w + 2 * x) * gC + ch;
    const float sum = f(g[b00]) + f(g[b00 + gC]) + f(g[b00 + (size_t)w * gC]) + f(g[b00 + (size_t)w * gC + gC]);
    out[i] = (f(act[i]) > 0.0f) ? (bf16)sum : (bf16)0.0f;
}

void launch_upsample_bwd(const bf16* g, int gC, const bf16* act, bf16* out, int n, int h, int w, int ca,
                         hipStream_t s) {
    const long long total = (long long)n * (h / 2) * (w / 2) * ca;
    hipLaunchKernelGGL(k_upsample_bwd, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, gC, act, out, n, h,
                       w, ca);
}

__global__ void k_pool_bwd_add(const bf16* gp, const bf16* gcat, int gcatC, int off, const bf16* act, bf16* out, int n,
                               int h, int w, int c) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * h * w * c;
    if (i >= total) return;
    const int ch = (int)(i % c);
    const long long q = i / c;
    const int x = (int)(q % w);
    const long long r = q / w;
    const int y = (int)(r % h), nn = (int)(r / h);
    const size_t pp = (((size_t)nn * (h >> 1) + (y >> 1)) * (w >> 1) + (x >> 1)) * c + ch;
    const float v = 0.25f * f(gp[pp]) + f(gcat[(size_t)q * gcatC + off + ch]);
    out[i] = (f(act[i]) > 0.0f) ? (bf16)v : (bf16)0.0f;
}

void launch_pool_bwd_add(const bf16* gp, const bf16* gcat, int gcatC, int off, const bf16* act, bf16* out, int n,
                         int h, int w, int c, hipStream_t s) {
    const long long total = (long long)n * h * w * c;
    hipLaunchKernelGGL(k_pool_bwd_add, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, gp, gcat, gcatC, off,
                       act, out, n, h, w, c);
}

// ---------------------------------------------------------------------------
// loss, optimiser, weight layouts
// ---------------------------------------------------------------------------
// Loss over the 3 output channels of every pixel, mean over pixels and
// channels (invCount = 1 / (3 P)).  mode 0: L1 in log space, |o - t|,
// gradient sign(o - t).  mode 1: relative MSE on linear radiance,
// (y - T)^2 / (yd^2 + 0.01) with y = e^o - 1, T = e^t - 1 and yd = max(y, 0)
// held constant (the evaluation's error measure), gradient
// 2 (y - T) e^o / (yd^2 + 0.01) clipped to [-4, 4].
__global__ __launch_bounds__(kBlock) void k_loss(const float* o, const float* t, long long pixels, float invCount,
                                                 int mode, bf16* go, int goC, float* lossAcc) {
    __shared__ float part[kBlock / 64];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (p < pixels) {
        for (int c = 0; c < goC; ++c) {
            float gv = 0.0f;
            if (c < kOutChannels) {
                const float oc = o[(size_t)p * 3 + c], tc = t[(size_t)p * 3 + c];
                if (mode == 0) {
                    const float d = oc - tc;
                    l += fabsf(d);
                    gv = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
                } else {
                    const float eo = expf(oc);
                    const float y = eo - 1.0f, T = expm1f(tc);
                    const float yd = fmaxf(y, 0.0f);
                    const float den = yd * yd + 0.01f;
                    const float d = y - T;
                    l += d * d / den;
                    gv = fminf(fmaxf(2.0f * d * eo / den, -4.0f), 4.0f);
                }
                gv *= invCount;
            }
            go[(size_t)p * goC + c] = (bf16)gv;
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < kBlock / 64; ++i) s += part[i];
        atomicAdd(lossAcc, s * invCount);
    }
}

void launch_loss(const float* o, const float* t, long long pixels, float invCount, int mode, bf16* go, int goC,
                 float* lossAcc, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3((unsigned)((pixels + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, o, t, pixels,
                       invCount, mode, go, goC, lossAcc);
}

// AdamW (decoupled weight decay), in the order w <- w (1 - lr wd);
// w <- w - lr mhat / (sqrt(vhat) + eps).
__global__ void k_adamw(AdamArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const float gr = a.g[i];
    const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * gr;
    const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * gr * gr;
    a.m[i] = m;
    a.v[i] = v;
    float w = a.w[i];
    if (a.decay) w = w * (1.0f - a.lr * a.wd);
    const float mh = m / a.bc1, vh = v / a.bc2;
    w = w - a.lr * mh / (sqrtf(vh) + a.eps);
    a.w[i] = w;
}

void launch_adamw(const AdamArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    hipLaunchKernelGGL(k_adamw, dim3((unsigned)((a.count + 255) / 256)), dim3(256), 0, s, a);
}

__global__ void k_weight_copies(const float* w, int coutPad, int cin, int coutT, bf16* w16, bf16* wt16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)coutPad * 9 * cin;
    if (i >= total) return;
    const int ci = (int)(i % cin);
    const int t = (int)((i / cin) % 9);
    const int co = (int)(i / ((long long)9 * cin));
    const bf16 v = (bf16)w[i];
    w16[i] = v;
    if (wt16 && co < coutT) wt16[((size_t)ci * 9 + (8 - t)) * coutT + co] = v;
}

void launch_weight_copies(const float* w, int coutPad, int cin, int coutT, bf16* w16, bf16* wt16, hipStream_t s) {
    const long long total = (long long)coutPad * 9 * cin;
    hipLaunchKernelGGL(k_weight_copies, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, w, coutPad, cin, coutT,
                       w16, wt16);
}

// ---------------------------------------------------------------------------
// render buffers <-> network tensors
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fireflyLum(float4 c) {
    return (0.2126f * c.x + 0.7152f * c.y) + 0.0722f * c.z;
}

// Image b of the `flips` images (1, 2 or 4) is the frame mirrored left-right
// when bit 0 of b is set and top-bottom when bit 1 is set, in the top-left
// w x h corner of its wp x hp tensor (zero padding right and below).  One
// thread per tensor position: it computes the features of the frame's pixel
// there once and writes them to that pixel's mirrored position in every image.
// One thread per pixel of the (padded) plane computes the features; the 64-B
// records of a wave's 64 pixels are then written through LDS so that each
// store instruction covers 16 whole records (1 KB) instead of a 16-B piece
// of 64 records.
__global__ __launch_bounds__(256) void k_features(const float4* rgb, const float4* alb, const float4* nrm, int w,
                                                  int h, int wp, int hp, int flips, bf16* x, bf16* rgb4) {
    static_assert(kInPad == 32 && kInChannels <= 16, "records of 4 16-byte pieces, the last two zero");
    __shared__ bf16x8 sRec[256 / 64][64][2];
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long plane = (long long)wp * hp;
    const bool valid = j < plane;
    const int px = valid ? (int)(j % wp) : 0, py = valid ? (int)(j / wp) : 0;
    float v[kInChannels] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const bool in = valid && px < w && py < h;
    if (in) {
        const size_t q = (size_t)py * w + px;
        float4 c = rgb[q];
        c.x = fmaxf(c.x, 0.0f);
        c.y = fmaxf(c.y, 0.0f);
        c.z = fmaxf(c.z, 0.0f);
        // firefly suppression: a pixel brighter than kFireflyRatio times the
        // brightest of its in-image 8 neighbours is scaled down to that bound
        const float lum = fireflyLum(c);
        float nb = 0.0f;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = px + dx, ny = py + dy;
                if ((dx | dy) == 0 || nx < 0 || ny < 0 || nx >= w || ny >= h) continue;
                float4 cn = rgb[(size_t)ny * w + nx];
                cn.x = fmaxf(cn.x, 0.0f);
                cn.y = fmaxf(cn.y, 0.0f);
                cn.z = fmaxf(cn.z, 0.0f);
                nb = fmaxf(nb, fireflyLum(cn));
            }
        const float bound = kFireflyRatio * nb;
        if (lum > bound) {
            const float k = bound / lum;
            c.x = c.x * k;
            c.y = c.y * k;
            c.z = c.z * k;
        }
        v[0] = log1pf(c.x);
        v[1] = log1pf(c.y);
        v[2] = log1pf(c.z);
        if (alb) {
            const float4 a = alb[q], n = nrm[q];
            v[3] = a.x; v[4] = a.y; v[5] = a.z;
            v[6] = n.x; v[7] = n.y; v[8] = n.z;
        }
    }
    bf16x8 out[kInPad / 8];
#pragma unroll
    for (int k = 0; k < kInPad / 8; ++k) out[k] = zero8();
#pragma unroll
    for (int c = 0; c < kInChannels; ++c) out[c / 8][c % 8] = (bf16)v[c];
    bf16x4 r4;
    r4[0] = out[0][0]; r4[1] = out[0][1]; r4[2] = out[0][2]; r4[3] = (bf16)0.0f;
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    sRec[wv][lane][0] = out[0];
    sRec[wv][lane][1] = out[1];
    __syncthreads();
    const long long jw = j - lane;   // the wave's first pixel
    const int q = lane & 3;          // this lane's 16-byte piece of a record
    for (int b = 0; b < flips; ++b) {
        // padding stays in place; frame pixels go to their mirrored position
        if (rgb4 && valid) {
            const int ox = (in && (b & 1)) ? w - 1 - px : px, oy = (in && (b & 2)) ? h - 1 - py : py;
            const size_t o = (size_t)b * (size_t)plane + (size_t)oy * wp + ox;
            *reinterpret_cast<bf16x4*>(rgb4 + o * 4) = r4;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = (lane >> 2) + 16 * k;   // record p of the wave, piece q
            const long long jp = jw + p;
            if (jp >= plane) continue;
            const int ppx = (int)(jp % wp), ppy = (int)(jp / wp);
            const bool pin = ppx < w && ppy < h;
            const int ox = (pin && (b & 1)) ? w - 1 - ppx : ppx, oy = (pin && (b & 2)) ? h - 1 - ppy : ppy;
            const size_t o = (size_t)b * (size_t)plane + (size_t)oy * wp + ox;
            *reinterpret_cast<bf16x8*>(x + o * kInPad + 8 * q) = q < 2 ? sRec[wv][p][q] : zero8();
        }
    }
}

void launch_features(const float4* rgb, const float4* alb, const float4* nrm, int w, int h, int wp, int hp,
                     int flips, bf16* x, bf16* rgb4, hipStream_t s) {
    const long long plane = (long long)wp * hp;
    hipLaunchKernelGGL(k_features, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, rgb, alb, nrm, w, h, wp,
                       hp, flips, x, rgb4);
}

__global__ void k_input_rgb(const bf16* x, long long pixels, bf16* rgb4) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    const bf16* src = x + (size_t)i * kInPad;
    bf16* r = rgb4 + (size_t)i * 4;
    r[0] = src[0]; r[1] = src[1]; r[2] = src[2]; r[3] = (bf16)0.0f;
}

void launch_input_rgb(const bf16* x, long long pixels, bf16* rgb4, hipStream_t s) {
    hipLaunchKernelGGL(k_input_rgb, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, s, x, pixels, rgb4);
}

// Mean of the `flips` outputs in log space, each mirrored back (image b as in
// k_features), summed in order b = 0, 1, ... so the result is deterministic.
__global__ void k_output(const float* o, int wp, int hp, int w, int h, int flips, float4* out, uint32_t* fb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)w * h) return;
    const int px = (int)(i % w), py = (int)(i / w);
    float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
    for (int b = 0; b < flips; ++b) {
        const int sx = (b & 1) ? w - 1 - px : px, sy = (b & 2) ? h - 1 - py : py;
        const float* src = o + (((size_t)b * hp + sy) * wp + sx) * 3;
        l0 += src[0];
        l1 += src[1];
        l2 += src[2];
    }
    const float inv = 1.0f / (float)flips;
    const float r = fmaxf(expm1f(l0 * inv), 0.0f), g = fmaxf(expm1f(l1 * inv), 0.0f), bl = fmaxf(expm1f(l2 * inv), 0.0f);
    out[i] = make_float4(r, g, bl, 0.0f);
    if (fb) fb[i] = packGamma(V3{r, g, bl});
}

void launch_output(const float* o, int wp, int hp, int w, int h, int flips, float4* out, uint32_t* fb,
                   hipStream_t s) {
    const long long total = (long long)w * h;
    hipLaunchKernelGGL(k_output, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, o, wp, hp, w, h, flips, out,
                       fb);
}

__global__ void k_log_target(const float4* rgb, long long pixels, float* t) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    const float4 c = rgb[i];
    t[3 * i] = log1pf(fmaxf(c.x, 0.0f));
    t[3 * i + 1] = log1pf(fmaxf(c.y, 0.0f));
    t[3 * i + 2] = log1pf(fmaxf(c.z, 0.0f));
}

void launch_log_target(const float4* rgb, long long pixels, float* t, hipStream_t s) {
    hipLaunchKernelGGL(k_log_target, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, s, rgb, pixels, t);
}

__global__ void k_fill_f32(float* p, long long n, float v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_f32(float* p, long long n, float v, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_copy_bytes(void* dst, const void* src, long long bytes, hipStream_t s) {
    const long long n = bytes / 4;   // callers copy whole 4-byte words
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, static_cast<uint32_t*>(dst),
                       static_cast<const uint32_t*>(src), n);
}

}  // namespace dn
}  // namespace bte
