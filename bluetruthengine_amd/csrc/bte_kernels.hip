// bte_kernels.hip — CDNA4 (gfx950) kernels of the wavefront path tracer.
//
// Pipeline per wavefront pass (frames [f0, f0+F) of every pixel in flight):
//   generate -> maxDepth x { extend, shade, connect[kind 0..K-1] } -> present
// All counts stay on the device: every kernel is persistent (grid sized to the
// machine) and pulls 64-item chunks with one atomic per wave, so the host never
// reads a counter back between stages.
//
// Reference counterparts: src/wavefront/ray_generate.cu:42-80 (generate),
// ray_extend.cu:49-72 + src/optix_programs.cu:110-148 (extend), ray_shade.cu:103-183
// (shade), ray_connect.cu:52-72 (connect), ray_present.cu:43-62 (present),
// src/optix_programs.cu:54-108 (megakernel), src/host/scene_upload.cu:428-459
// (env-map CDF).
#include <algorithm>

#include "bte_kernels.h"
#include "bte_device.h"

namespace bte {

constexpr int kBlock = 256;
constexpr int kLdsStack = 32;                 // LDS stack entries per lane (megakernel, trace_rays)
constexpr int kMinLdsStack = 8;               // smallest LDS stack of any traversal kernel variant
constexpr int kOvfStack = kStackSize - kMinLdsStack;   // global overflow entries per thread
// path word st.y: path id (bits 0-29) | AOV recorded (bit 30) | specular (bit 31)
constexpr uint32_t kSpecBit = 0x80000000u;
constexpr uint32_t kAovBit = 0x40000000u;
constexpr uint32_t kPidMask = 0x3fffffffu;
constexpr uint32_t kFinalBit = 0x80000000u;

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
BTE_DEV int lane_id() { return (int)(threadIdx.x & 63u); }

// Block-aggregated append: the waves' counts meet in LDS and one lane adds
// the counter's total, so a counter that every wave of the grid appends to
// sees one atomic per block instead of one per wave (same-address device
// atomics serialise).  MULTI: waves may pass different counters (the
// per-scene queues of a batch), and waves with the same counter share one
// atomic.  sh holds 2 ints per wave, ctrs one pointer per wave (MULTI only).
// Every wave of the block must call it the same number of times;
// consecutive calls may share sh and ctrs.
template <bool MULTI>
BTE_DEV int block_append(int* ctr, bool pred, int* sh, int** ctrs) {
    constexpr int kWaves = kBlock / 64;
    const unsigned long long mask = __ballot(pred);
    const int w = (int)(threadIdx.x >> 6);
    if (lane_id() == 0) {
        sh[w] = __popcll(mask);
        if (MULTI) ctrs[w] = ctr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if constexpr (!MULTI) {
            int total = 0;
            for (int i = 0; i < kWaves; ++i) {
                const int c = sh[i];
                sh[kWaves + i] = total;
                total += c;
            }
            const int base = total ? atomicAdd(ctr, total) : 0;
            for (int i = 0; i < kWaves; ++i) sh[kWaves + i] += base;
        } else {
            // waves in order; each counter's total goes to the first wave using it
            int first[kWaves];
            for (int i = 0; i < kWaves; ++i) {
                first[i] = i;
                for (int j = 0; j < i; ++j)
                    if (first[i] == i && ctrs[j] == ctrs[i]) first[i] = j;
            }
            int run[kWaves];
            for (int i = 0; i < kWaves; ++i) run[i] = 0;
            for (int i = 0; i < kWaves; ++i) {
                const int c = sh[i];
                sh[kWaves + i] = run[first[i]];
                run[first[i]] += c;
            }
            for (int i = 0; i < kWaves; ++i)
                if (first[i] == i) run[i] = run[i] ? atomicAdd(ctrs[i], run[i]) : 0;
            for (int i = 0; i < kWaves; ++i) sh[kWaves + i] += run[first[i]];
        }
    }
    __syncthreads();
    const unsigned lo = (unsigned)(mask & 0xffffffffull), hi = (unsigned)(mask >> 32);
    const int prefix = (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    return pred ? sh[kWaves + w] + prefix : -1;
}

// LDS of a block's appends (block_append)
struct BlockAppendLds {
    int n[2 * (kBlock / 64)];
    int* ctrs[kBlock / 64];
};

BTE_DEV void wave_add_u64(unsigned long long* ctr, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane_id() == 0 && x) atomicAdd(ctr, x);
}

BTE_DEV int wave_grab(int* ctr, int n = 64) {
    int base = 0;
    if (lane_id() == 0) base = atomicAdd(ctr, n);
    return __shfl(base, 0, 64);
}

// Work chunks of a persistent wave: the first one is fixed by the wave's index
// in the grid, later ones come from the counter (offset past the fixed ones).
// Device-wide atomics on one address serialise, so a wave whose fixed chunk
// is already past the end leaves without touching the counter: a bounce with
// little work costs no atomic storm.
// Later chunks are dealt out by one counter per XCD (group = kWorkXcds
// counters): counter x hands out chunks fixedEnd + (i * nx + x) * n, so every
// chunk goes to exactly one wave whatever the block-to-XCD placement.
BTE_DEV int first_chunk(int n) { return (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * n; }
BTE_DEV int next_chunk(int* group, int n, int count) {
    const int fixedEnd = (int)(gridDim.x * (blockDim.x >> 6)) * n;
    if (fixedEnd >= count) return count;
    const int nx = min((int)gridDim.x, kWorkXcds);
    const int x = (int)blockIdx.x % nx;
    const int i = wave_grab(group + x * kWorkPad, 1);
    return fixedEnd + (i * nx + x) * n;
}
BTE_DEV int* work_group(const WaveBuffers& wb, int bounce, int g) {
    return wb.work + (size_t)bounce * kWorkStride + (size_t)g * (kWorkXcds * kWorkPad);
}

// ---------------------------------------------------------------------------
// multi-scene batches
// A batch keeps one DevScene per scene in device memory.  Every queue keeps
// each scene's entries at the start of that scene's own slot range and is
// walked in compact rows (scene after scene, empty ranges skipped), so a work
// chunk never spans two scenes: the scene index is wave-uniform and its
// descriptor arrives through scalar loads, as a kernel argument would.
// ---------------------------------------------------------------------------
// Per-scene counts and compact row starts of the batch's queues.
BTE_DEV int* segCounts(const WaveBuffers& wb, int bounce, int q) {
    return wb.segCount + ((size_t)bounce * kSegQueues + (size_t)q) * (size_t)wb.sceneCount;
}
BTE_DEV int* segRowStart(const WaveBuffers& wb, int bounce, int q) {
    return wb.segRows + ((size_t)bounce * kSegQueues + (size_t)q) * (size_t)(wb.sceneCount + 1);
}
// Scene of compact row r (r wave-uniform): the last s with start[s] <= r, so
// scenes without rows are skipped.
BTE_DEV int segSceneOfRow(const int* start, int sceneCount, int r) {
    int lo = 0, hi = sceneCount - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}
// A batch work chunk (compact slot index c, wave-uniform): its scene s, the
// queue position of its first entry and the number of entries left in s.
struct SegChunk {
    int s, first, left;
};
BTE_DEV SegChunk segChunk(const WaveBuffers& wb, int bounce, int q, int c) {
    const int* start = segRowStart(wb, bounce, q);
    const int r = c >> 6;
    SegChunk k;
    k.s = segSceneOfRow(start, wb.sceneCount, r);
    const int local = (r - start[k.s]) * 64;
    k.first = (int)wb.slots[k.s].pathBase + local;
    k.left = segCounts(wb, bounce, q)[k.s] - local;
    return k;
}
BTE_DEV int segSlots(const WaveBuffers& wb, int bounce, int q) {
    return segRowStart(wb, bounce, q)[wb.sceneCount] * 64;
}

// A pointer read from memory is generic to the compiler; rebuilt from its
// integer value as a global-space pointer, the loads through it become global
// (not flat) loads.
template <class T>
BTE_DEV const T* globalPtr(const T* p) {
    return (const T*)(const __attribute__((address_space(1))) T*)(uintptr_t)p;
}

// scenes must be a __restrict__ kernel argument and s wave-uniform.
BTE_DEV DevScene fetchScene(const DevScene* __restrict__ scenes, int s) {
    DevScene sc = scenes[s];
    sc.nodes = globalPtr(sc.nodes);
    sc.nodes4 = globalPtr(sc.nodes4);
    sc.nodes8 = globalPtr(sc.nodes8);
    sc.leafTris = globalPtr(sc.leafTris);
    sc.tris = globalPtr(sc.tris);
    sc.triBsdf = globalPtr(sc.triBsdf);
    sc.bsdfs = globalPtr(sc.bsdfs);
    sc.triEmission = globalPtr(sc.triEmission);
    sc.triFlags = globalPtr(sc.triFlags);
    sc.emitterOfTri = globalPtr(sc.emitterOfTri);
    sc.emitters = globalPtr(sc.emitters);
    sc.emTriIdx = globalPtr(sc.emTriIdx);
    sc.emTriCdf = globalPtr(sc.emTriCdf);
    sc.sceneCdf = globalPtr(sc.sceneCdf);
    sc.spots = globalPtr(sc.spots);
    sc.env = globalPtr(sc.env);
    sc.envMarg = globalPtr(sc.envMarg);
    sc.envCond = globalPtr(sc.envCond);
    sc.envMargGuide = globalPtr(sc.envMargGuide);
    sc.envCondGuide = globalPtr(sc.envCondGuide);
    sc.triMat = globalPtr(sc.triMat);
    sc.texels = globalPtr(sc.texels);
    sc.texInfo = globalPtr(sc.texInfo);
    sc.mrTexels = globalPtr(sc.mrTexels);
    sc.mrInfo = globalPtr(sc.mrInfo);
    return sc;
}

// ---------------------------------------------------------------------------
// BVH2 traversal (replaces optixTrace: rt_optix_shading.cuh:57-95)
// Watertight ray/triangle test of Woop, Benthin & Wald (JCGT 2013); ties on t
// resolve to the lower triangle index so the result is independent of the
// tree.  Stack: kLdsStack entries per lane in LDS (stride kBlock, conflict
// free), the rest in a per-thread global overflow slice.
// ---------------------------------------------------------------------------
struct WtRay {
    V3 o, d;
    float inv[3];
    float Sx, Sy, Sz;
    int kx, ky, kz;
    bool valid;
};
BTE_DEV float comp(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

BTE_DEV WtRay makeWtRay(V3 o, V3 d) {
    WtRay r;
    r.o = o;
    r.d = d;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    r.valid = (ax > 0.0f || ay > 0.0f || az > 0.0f);
    int kz = 0;
    float best = ax;
    if (ay > best) { kz = 1; best = ay; }
    if (az > best) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    const float dz = comp(d, kz);
    if (dz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.inv[0] = 1.0f / d.x;
    r.inv[1] = 1.0f / d.y;
    r.inv[2] = 1.0f / d.z;
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = kz == 0 ? r.inv[0] : (kz == 1 ? r.inv[1] : r.inv[2]);   // = 1 / dz, one division fewer
    return r;
}

// Line interval of the box of three points, in the arithmetic of boxEnter;
// false when it is empty.
BTE_DEV bool triSpan(const WtRay& r, float4 a, float4 b, float4 c, float& lo, float& hi) {
    const float x0 = (fminf(fminf(a.x, b.x), c.x) - r.o.x) * r.inv[0];
    const float x1 = (fmaxf(fmaxf(a.x, b.x), c.x) - r.o.x) * r.inv[0];
    const float y0 = (fminf(fminf(a.y, b.y), c.y) - r.o.y) * r.inv[1];
    const float y1 = (fmaxf(fmaxf(a.y, b.y), c.y) - r.o.y) * r.inv[1];
    const float z0 = (fminf(fminf(a.z, b.z), c.z) - r.o.z) * r.inv[2];
    const float z1 = (fmaxf(fmaxf(a.z, b.z), c.z) - r.o.z) * r.inv[2];
    lo = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
    hi = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
    return lo <= hi;
}

// Returns true with (t, u, v) when the ray hits within [tmin, tmax].  The
// distance T/det of a ray nearly parallel to a large triangle is badly
// conditioned, so it is clamped to the line interval of the triangle's own
// box (a triangle whose interval is empty is missed).  Every BVH box holding
// the triangle then admits that distance in boxEnter's arithmetic (rounding
// is monotone), so a traversal finds exactly what a loop over all triangles
// finds, whatever the tree.
BTE_DEV bool hitTriangle(const WtRay& r, float4 a, float4 b, float4 c, float tmin, float tmax,
                         float& tOut, float& uOut, float& vOut) {
    const V3 A = sub3(xyz(a), r.o);
    const V3 B = sub3(xyz(b), r.o);
    const V3 C = sub3(xyz(c), r.o);
    const float Az_ = comp(A, r.kz), Bz_ = comp(B, r.kz), Cz_ = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Az_;
    const float Ay = comp(A, r.ky) - r.Sy * Az_;
    const float Bx = comp(B, r.kx) - r.Sx * Bz_;
    const float By = comp(B, r.ky) - r.Sy * Bz_;
    const float Cx = comp(C, r.kx) - r.Sx * Cz_;
    const float Cy = comp(C, r.ky) - r.Sy * Cz_;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Az_;
    const float Bz = r.Sz * Bz_;
    const float Cz = r.Sz * Cz_;
    const float T = U * Az + V * Bz + W * Cz;
    float lo, hi;
    if (!triSpan(r, a, b, c, lo, hi)) return false;
    const float t = fminf(fmaxf(T / det, lo), hi);
    if (!(t >= tmin && t <= tmax)) return false;
    tOut = t;
    uOut = V / det;
    vOut = W / det;
    return true;
}

// Slab interval of one child box: entry distance returned, far distance
// (limited by the query's current tmax / best hit) in tfar.  No widening is
// needed: hitTriangle clamps a hit to its triangle's own interval, which any
// enclosing box's computed interval contains.
BTE_DEV float boxEnter(const WtRay& r, float lx, float hx, float ly, float hy, float lz, float hz, float tmin,
                       float tmax, float& tfar) {
    const float ax = (lx - r.o.x) * r.inv[0], bx = (hx - r.o.x) * r.inv[0];
    const float ay = (ly - r.o.y) * r.inv[1], by = (hy - r.o.y) * r.inv[1];
    const float az = (lz - r.o.z) * r.inv[2], bz = (hz - r.o.z) * r.inv[2];
    const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    const float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    tfar = fminf(t1, tmax);
    return t0;
}

// Traversal stack: the first L entries of each lane live in LDS (stride kBlock,
// conflict free), deeper entries in the lane's global overflow slice.
template <int L>
struct StackRef {
    int* lds;      // &lds[threadIdx.x]
    int* ovf;      // per-thread global slice (kOvfStack entries)
};
template <int L>
BTE_DEV void push(const StackRef<L>& s, int& sp, int v) {
    if (sp < L) s.lds[sp * kBlock] = v;
    else s.ovf[sp - L] = v;
    ++sp;
}
template <int L>
BTE_DEV int pop(const StackRef<L>& s, int& sp) {
    --sp;
    // an LDS read for every lane (clamped slot), the overflow slice only
    // where the stack is that deep (a volatile read, so that the two loads
    // are not merged into one flat load, whose wait covers every
    // outstanding memory access of the wave)
    int v = s.lds[min(sp, L - 1) * kBlock];
    if (sp >= L) v = *(const volatile int*)(s.ovf + (sp - L));
    return v;
}

// ANY = false: closest hit, returns triangle id (or -1) with t,u,v.
// ANY = true : returns 1 if any triangle is hit within [tmin, tmax], else 0.
template <bool ANY, bool COUNT = false, int L = kLdsStack>
BTE_DEV int traverse(const DevScene& sc, V3 o, V3 d, float tmin, float tmax, const StackRef<L>& stk, float& tOut,
                     float& uOut, float& vOut, uint32_t* nodeCount = nullptr, uint32_t* triCount = nullptr) {
    const WtRay r = makeWtRay(o, d);
    if (!r.valid || sc.triCount == 0) return ANY ? 0 : -1;
    float best = tmax;
    int bestTri = -1;
    int node = 0;
    int sp = 0;
    for (;;) {
        if (node >= 0) {
            if (COUNT) ++*nodeCount;
            const float4* np = sc.nodes + 4 * (size_t)node;
            const float4 n0 = np[0], n1 = np[1], n2 = np[2];
            const int4 ch = *reinterpret_cast<const int4*>(np + 3);
            float f0, f1;
            const float e0 = boxEnter(r, n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, tmin, best, f0);
            const float e1 = boxEnter(r, n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, tmin, best, f1);
            const bool h0 = e0 <= f0, h1 = e1 <= f1;
            if (h0 && h1) {
                const bool first0 = e0 <= e1;
                node = first0 ? ch.x : ch.y;
                push(stk, sp, first0 ? ch.y : ch.x);
                continue;
            }
            if (h0) { node = ch.x; continue; }
            if (h1) { node = ch.y; continue; }
        } else {
            const int code = ~node;
            const int start = code >> kLeafCountBits;
            const int cnt = (code & (kMaxLeafTris - 1)) + 1;
            for (int i = 0; i < cnt; ++i) {
                const float4* tp = sc.leafTris + 3 * (size_t)(start + i);
                const float4 a = tp[0], b = tp[1], c = tp[2];
                if (COUNT) ++*triCount;
                float t, u, v;
                if (hitTriangle(r, a, b, c, tmin, best, t, u, v)) {
                    if (ANY) return 1;
                    const int id = __float_as_int(a.w);
                    if (t < best || bestTri < 0 || id < bestTri) {
                        best = t; bestTri = id; tOut = t; uOut = u; vOut = v;
                    }
                }
            }
        }
        if (sp == 0) break;
        node = pop(stk, sp);
    }
    return ANY ? 0 : bestTri;
}

// Compare-exchange by key, written as selects: a branch per pair costs the
// wave an exec-mask save/restore and the moves of both arms.
BTE_DEV void sortPair(float& ka, int& ca, float& kb, int& cb) {
    const bool sw = kb < ka;
    const float lo = sw ? kb : ka, hi = sw ? ka : kb;
    const int clo = sw ? cb : ca, chi = sw ? ca : cb;
    ka = lo; kb = hi;
    ca = clo; cb = chi;
}

// Pushes c3, c2, c1 in this order (c1 on top), each only when its flag is
// set.  While all three slots lie in LDS the stores are unconditional: the
// slot of an unset child is overwritten by the next store or left above the
// top.  Same stack contents as three conditional pushes.
template <int L>
BTE_DEV void push3(const StackRef<L>& s, int& sp, int c1, bool v1, int c2, bool v2, int c3, bool v3) {
    if (sp + 2 < L) {
        int p = sp;
        s.lds[p * kBlock] = c3;
        p += v3 ? 1 : 0;
        s.lds[p * kBlock] = c2;
        p += v2 ? 1 : 0;
        s.lds[p * kBlock] = c1;
        p += v1 ? 1 : 0;
        sp = p;
    } else {
        if (v3) push(s, sp, c3);
        if (v2) push(s, sp, c2);
        if (v1) push(s, sp, c1);
    }
}

// BVH4 traversal (same contract as traverse): for closest hits the hit
// children of a node are ordered by entry distance, the nearest visited next
// and the others pushed farthest first; any-hit queries skip the ordering.
template <bool ANY, bool COUNT = false, int L = kLdsStack>
BTE_DEV int traverse4(const DevScene& sc, V3 o, V3 d, float tmin, float tmax, const StackRef<L>& stk, float& tOut,
                      float& uOut, float& vOut, uint32_t* nodeCount = nullptr, uint32_t* triCount = nullptr) {
    const WtRay r = makeWtRay(o, d);
    if (!r.valid || sc.triCount == 0) return ANY ? 0 : -1;
    constexpr float kNone = __builtin_huge_valf();
    float best = tmax;
    int bestTri = -1;
    int node = 0;
    int sp = 0;
    for (;;) {
        if (node >= 0) {
            if (COUNT) ++*nodeCount;
            const float4* np = sc.nodes4 + 8 * (size_t)node;
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const int4 ch = *reinterpret_cast<const int4*>(np + 6);
            float f0, f1, f2, f3;
            const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, tmin, best, f0);
            const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, tmin, best, f1);
            const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, tmin, best, f2);
            const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, tmin, best, f3);
            float k0 = e0 <= f0 ? e0 : kNone, k1 = e1 <= f1 ? e1 : kNone;
            float k2 = e2 <= f2 ? e2 : kNone, k3 = e3 <= f3 ? e3 : kNone;
            int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
            if constexpr (!ANY) {
                sortPair(k0, c0, k1, c1);
                sortPair(k2, c2, k3, c3);
                sortPair(k0, c0, k2, c2);
                sortPair(k1, c1, k3, c3);
                sortPair(k1, c1, k2, c2);
            } else {
                // any hit ends the search: order is irrelevant, only move a hit child to slot 0
                if (k0 == kNone) {
                    if (k1 != kNone) { k0 = k1; c0 = c1; k1 = kNone; }
                    else if (k2 != kNone) { k0 = k2; c0 = c2; k2 = kNone; }
                    else if (k3 != kNone) { k0 = k3; c0 = c3; k3 = kNone; }
                }
            }
            if (k0 != kNone) {
                push3(stk, sp, c1, k1 != kNone, c2, k2 != kNone, c3, k3 != kNone);
                node = c0;
                continue;
            }
        } else {
            const int code = ~node;
            const int start = code >> kLeafCountBits;
            const int cnt = (code & (kMaxLeafTris - 1)) + 1;
            for (int i = 0; i < cnt; ++i) {
                const float4* tp = sc.leafTris + 3 * (size_t)(start + i);
                const float4 a = tp[0], b = tp[1], c = tp[2];
                if (COUNT) ++*triCount;
                float t, u, v;
                if (hitTriangle(r, a, b, c, tmin, best, t, u, v)) {
                    if (ANY) return 1;
                    const int id = __float_as_int(a.w);
                    if (t < best || bestTri < 0 || id < bestTri) {
                        best = t; bestTri = id; tOut = t; uOut = u; vOut = v;
                    }
                }
            }
        }
        if (sp == 0) break;
        node = pop(stk, sp);
    }
    return ANY ? 0 : bestTri;
}

template <int W, bool ANY, bool COUNT, int L>
BTE_DEV int traverseW(const DevScene& sc, V3 o, V3 d, float tmin, float tmax, const StackRef<L>& stk, float& t,
                      float& u, float& v, uint32_t* nodeCount, uint32_t* triCount) {
    if constexpr (W == 4) return traverse4<ANY, COUNT, L>(sc, o, d, tmin, tmax, stk, t, u, v, nodeCount, triCount);
    else return traverse<ANY, COUNT, L>(sc, o, d, tmin, tmax, stk, t, u, v, nodeCount, triCount);
}

// ---------------------------------------------------------------------------
// Scene helpers
// ---------------------------------------------------------------------------
BTE_DEV Bsdf loadBsdfId(const DevScene& sc, int id) {
    const float4* p = sc.bsdfs + 4 * (size_t)id;
    Bsdf b;
    b.p0 = p[0]; b.p1 = p[1]; b.p2 = p[2]; b.p3 = p[3];
    b.type = __float_as_int(b.p3.w);
    return b;
}
BTE_DEV Bsdf loadBsdf(const DevScene& sc, int tri) {            // rt_material.cuh:15-24
    return loadBsdfId(sc, sc.triBsdf[tri]);
}

// RGBA8 texture lookup at normalized (u, v): bilinear, wrap addressing in both
// axes, texel centres at (i + 0.5) / size, bytes read as byte / 255 (the
// texture unit's normalized-float mode, evaluated in exact binary32).
BTE_DEV float4 texBilinear(const uint32_t* texels, int4 info, float u, float v) {
    const int W = info.y, H = info.z;
    u = u - floorf(u);
    v = v - floorf(v);
    if (!(u == u)) u = 0.0f;
    if (!(v == v)) v = 0.0f;
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    int ix0 = (int)fx, iy0 = (int)fy;
    if (ix0 < 0) ix0 += W;
    if (iy0 < 0) iy0 += H;
    if (ix0 >= W) ix0 -= W;
    if (iy0 >= H) iy0 -= H;
    const int ix1 = ix0 + 1 == W ? 0 : ix0 + 1, iy1 = iy0 + 1 == H ? 0 : iy0 + 1;
    const uint32_t* t = texels + info.x;
    const uint32_t t00 = t[(size_t)iy0 * W + ix0], t10 = t[(size_t)iy0 * W + ix1];
    const uint32_t t01 = t[(size_t)iy1 * W + ix0], t11 = t[(size_t)iy1 * W + ix1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    float r[4];
    for (int c = 0; c < 4; ++c) {
        const float c00 = (float)((t00 >> (8 * c)) & 0xffu) / 255.0f, c10 = (float)((t10 >> (8 * c)) & 0xffu) / 255.0f;
        const float c01 = (float)((t01 >> (8 * c)) & 0xffu) / 255.0f, c11 = (float)((t11 >> (8 * c)) & 0xffu) / 255.0f;
        r[c] = c00 * w00 + c10 * w10 + c01 * w01 + c11 * w11;
    }
    return make_float4(r[0], r[1], r[2], r[3]);
}

// Texture override of a hit's BSDF (rt_material.cuh:26-56): the albedo map
// replaces p0.xyz (microfacet ks recomputed from it), a metallic-roughness
// map scales a Disney BSDF's roughness (p0.w, by G) and metallic (p1.x, by B).
// (u, v) are the hit's barycentrics; the texture coordinate is interpolated
// from the triangle's (src/optix_programs.cu:137-140).
BTE_DEV void applyTextures(const DevScene& sc, int tri, float bu, float bv, Bsdf& b) {
    if (!sc.triMat || sc.texCount <= 0) return;
    const int m = sc.triMat[tri];
    if (m < 0 || m >= sc.texCount) return;
    const float* tp = reinterpret_cast<const float*>(sc.tris + tri) + 18;
    const float w0 = 1.0f - bu - bv;
    const float tu = w0 * tp[0] + bu * tp[2] + bv * tp[4];
    const float tv = w0 * tp[1] + bu * tp[3] + bv * tp[5];
    const int4 ti = sc.texInfo[m];
    if (ti.y > 0) {
        const float4 s = texBilinear(sc.texels, ti, tu, tv);
        b.p0.x = s.x;
        b.p0.y = s.y;
        b.p0.z = s.z;
        if (b.type == BTE_BSDF_MICROFACET) b.p1.w = 1.0f - fmaxf(b.p0.x, fmaxf(b.p0.y, b.p0.z));
    }
    if (m < sc.mrCount && b.type == BTE_BSDF_DISNEY) {
        const int4 mi = sc.mrInfo[m];
        if (mi.y > 0) {
            const float4 mr = texBilinear(sc.mrTexels, mi, tu, tv);
            b.p0.w = b.p0.w * mr.y;
            b.p1.x = b.p1.x * mr.z;
        }
    }
}

// Hit attributes — src/optix_programs.cu:127-141 (hit point O + tD, normalized
// barycentric shading normal, not face-forwarded).
struct Its { V3 p, n; int tri; };
BTE_DEV Its makeIts(const DevScene& sc, V3 o, V3 d, float4 h) {
    Its its;
    its.tri = __float_as_int(h.x);
    const float t = h.y, u = h.z, v = h.w;
    its.p = {o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
    const float* tn = reinterpret_cast<const float*>(sc.tris + its.tri) + 9;
    const float w0 = 1.0f - u - v;
    its.n = normalize3({w0 * tn[0] + u * tn[3] + v * tn[6], w0 * tn[1] + u * tn[4] + v * tn[7],
                        w0 * tn[2] + u * tn[5] + v * tn[8]});
    return its;
}

// Env map — rt_env_map.cuh:36-199 with exact binary32 bilinear filtering.
BTE_DEV V3 envTexel(const DevScene& sc, int ix, int iy) {
    const float4 t = sc.env[(size_t)iy * (size_t)sc.envW + (size_t)ix];
    return {t.x, t.y, t.z};
}
// The 2x2 texels of a bilinear environment lookup at (u, v): corners ix0 / ix1
// (wrapped) by iy0 / iy1 (clamped), weights a / b, and the footprint record
// (iyb + 1) * W + ix0 that holds them (iyb = floor(v H - 0.5) in [-1, H-1]).
struct EnvFoot {
    int ix0, ix1, iy0, iy1;
    float a, b;
    size_t rec;
};
BTE_DEV EnvFoot envFoot(const DevScene& sc, float u, float v) {
    const int W = sc.envW, H = sc.envH;
    float x = u * (float)W - 0.5f;
    float y = v * (float)H - 0.5f;
    if (!(x == x)) x = 0.0f;
    if (!(y == y)) y = 0.0f;
    const float fx = floorf(x), fy = floorf(y);
    EnvFoot f;
    f.a = x - fx;
    f.b = y - fy;
    const int ix = (int)fx, iy = (int)fy;
    int ix0 = ix % W; if (ix0 < 0) ix0 += W;
    int ix1 = ix0 + 1; if (ix1 == W) ix1 = 0;
    f.ix0 = ix0;
    f.ix1 = ix1;
    f.iy0 = min(max(iy, 0), H - 1);
    f.iy1 = min(max(iy + 1, 0), H - 1);
    const int iyb = min(max(iy, -1), H - 1);   // iy0 and iy1 depend on iy only through this
    f.rec = (size_t)(iyb + 1) * (size_t)W + (size_t)ix0;
    return f;
}
// Bilinear texel mix of a footprint record (r0..r2: the corners' RGB).
BTE_DEV V3 envMix(const EnvFoot& f, const float4& r0, const float4& r1, const float4& r2) {
    const float a = f.a, b = f.b;
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    const V3 t00 = {r0.x, r0.y, r0.z}, t10 = {r0.w, r1.x, r1.y}, t01 = {r1.z, r1.w, r2.x}, t11 = {r2.y, r2.z, r2.w};
    return {w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
            w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
            w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z};
}
BTE_DEV V3 envBilinear(const DevScene& sc, float u, float v) {
    const EnvFoot f = envFoot(sc, u, v);
    const float4* r = sc.envRec + 4 * f.rec;
    return envMix(f, r[0], r[1], r[2]);
}
BTE_DEV void dirToEquirect(V3 dir, float rot, float& u, float& v) {
    float phi = datan2(dir.z, dir.x) + rot;
    const float theta = dacos(fminf(fmaxf(dir.y, -1.0f), 1.0f));
    phi = fmodf(phi + m::PI, 2.0f * m::PI);
    if (phi < 0.0f) phi += 2.0f * m::PI;
    phi -= m::PI;
    u = (phi + m::PI) / (2.0f * m::PI);
    v = theta / m::PI;
}
BTE_DEV V3 envEval(const DevScene& sc, V3 dir) {
    float u, v;
    dirToEquirect(dir, sc.envRot, u, v);
    const V3 t = envBilinear(sc, u, v);
    const float k = sc.envIntensity;
    return {t.x * k, t.y * k, t.z * k};
}
BTE_DEV int cdfSearch(const float* cdf, int count, float x) {
    int lo = 0, hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}
// Same result as cdfSearch on a non-decreasing CDF (the index i with
// cdf[i] <= x < cdf[i+1] is unique): a guide table gives a bracket
// [guide[g], guide[g+1]] for g = floor(u*G), which is checked against x before
// the binary search runs inside it (full range if the check fails).
BTE_DEV int cdfSearchGuided(const float* cdf, int count, float x, const int* guide, int G, float u) {
    if (!guide) return cdfSearch(cdf, count, x);
    const int g = min(max((int)fminf(fmaxf(u * (float)G, 0.0f), (float)(G - 1)), 0), G - 1);
    int lo = guide[g];
    int hi = guide[g + 1] + 1;
    if (!(cdf[lo] <= x)) lo = 0;
    if (hi < count && !(cdf[hi] > x)) hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}
// The environment pdf of texel (ix, iy) before its division by
// 2 pi^2 sin(theta): ((pdfRow * pdfCol) * W) * H, or 0 where the pdf is 0.
BTE_DEV float envPdfFactor(const float* marg, const float* cond, int W, int H, float total, int ix, int iy) {
    if (total <= 0.0f) return 0.0f;
    const float ms = marg[H];
    const float rw = marg[iy + 1] - marg[iy];
    if (ms <= 0.0f || rw <= 0.0f) return 0.0f;
    const float pdfRow = rw / ms;
    const float* row = cond + (size_t)iy * (size_t)(W + 1);
    const float rs = row[W];
    const float cw = row[ix + 1] - row[ix];
    if (rs <= 0.0f || cw <= 0.0f) return 0.0f;
    const float pdfCol = cw / rs;
    return pdfRow * pdfCol * (float)W * (float)H;
}
// The pdf texel of (u, v), nearest rather than bilinear.
BTE_DEV void envPdfTexel(const DevScene& sc, float u, float v, int& ix, int& iy) {
    const int W = sc.envW, H = sc.envH;
    const float fx = u * (float)W;
    const float fy = v * (float)H;
    ix = min(max((int)fminf(fmaxf(fx, 0.0f), (float)W), 0), W - 1);
    iy = min(max((int)fminf(fmaxf(fy, 0.0f), (float)H), 0), H - 1);
}
BTE_DEV float envPdfFromFactor(float q, float v) {
    const float theta = v * m::PI;
    const float st = fmaxf(dsin(theta), 1e-6f);
    return q / (2.0f * m::PI * m::PI * st);
}
BTE_DEV float envPdfUV(const DevScene& sc, float u, float v) {
    int ix, iy;
    envPdfTexel(sc, u, v, ix, iy);
    return envPdfFromFactor(envPdfFactor(sc.envMarg, sc.envCond, sc.envW, sc.envH, sc.envTotal, ix, iy), v);
}

// ---------------------------------------------------------------------------
// Direct lighting — rt_direct_lighting.cuh, rt_emitter_sampling.cuh
// ---------------------------------------------------------------------------
struct NeeRec { V3 d; float tMax; V3 c; bool ok; };

BTE_DEV int cdfReuse(const float* cdf, int count, float& u, float& pdf) {   // :33-65
    const float sum = cdf[count];
    if (sum <= 0.0f || count <= 0) { pdf = 0.0f; u = 0.0f; return 0; }
    const float x = u * sum;
    int lo = 0, hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid;
    }
    const float c0 = cdf[lo], c1 = cdf[lo + 1];
    const float den = c1 - c0;
    if (den <= 0.0f) { u = 0.0f; pdf = 0.0f; return lo; }
    u = (x - c0) / den;
    pdf = den / sum;
    return lo;
}
BTE_DEV float areaToSolidAngle(float pdfArea, V3 hitPoint, V3 origin, V3 hitNormal, V3 dirToLight) {
    const V3 diff = sub3(hitPoint, origin);
    const float dist2 = dot3(diff, diff);
    const V3 negL = {-dirToLight.x, -dirToLight.y, -dirToLight.z};
    const float den = fmaxf(fabsf(dot3(hitNormal, negL)), 1e-3f);
    return pdfArea * dist2 / den;
}

// Emitter hit with BSDF-side MIS (:60-99).  rad += Le * throughput * w.
BTE_DEV void emitterHit(const DevScene& sc, V3 o, V3 d, bool spec, float brdfPdf, V3 thr, const Its& its, V3& rad) {
    V3 Le = {0, 0, 0};
    if (sc.triEmission) {
        const float* e = sc.triEmission + 3 * (size_t)its.tri;
        Le = {e[0], e[1], e[2]};
    }
    float w = 1.0f;
    if (!spec) {
        const int eIdx = sc.emitterOfTri ? sc.emitterOfTri[its.tri] : -1;
        if (eIdx >= 0) {
            const float areaSum = sc.emitters[eIdx].areaSum;
            const float areaPdf = (areaSum <= 0.0f) ? 0.0f : 1.0f / areaSum;
            float sel = 0.0f;
            const float sum = sc.sceneCdf[sc.emitterCount];
            if (sum > 0.0f) sel = (sc.sceneCdf[eIdx + 1] - sc.sceneCdf[eIdx]) / sum;
            const float lp = areaToSolidAngle(areaPdf * sel, its.p, o, its.n, d);
            const float den = brdfPdf + lp;
            w = (den <= 0.0f) ? 0.0f : brdfPdf / den;
        }
    }
    if (dot3(its.n, V3{-d.x, -d.y, -d.z}) <= 0.0f) Le = {0, 0, 0};   // checkRadiance (:112-125)
    rad = add3(rad, mul3(mul3(Le, thr), w));
}

// BSDF dispatch specialised by a compile-time type (T = 0: runtime dispatch on b.type).
template <int T>
BTE_DEV V3 bEval(const Bsdf& b, const BRec& r) {
    if constexpr (T == BTE_BSDF_DIFFUSE) return evalDiffuse(b, r);
    else if constexpr (T == BTE_BSDF_MICROFACET) return evalMicrofacet(b, r);
    else if constexpr (T == BTE_BSDF_DISNEY) return evalDisney(b, r);
    else if constexpr (T == BTE_BSDF_DIELECTRIC || T == BTE_BSDF_MIRROR) return V3{0, 0, 0};
    else return bsdfEval(b, r);
}
template <int T>
BTE_DEV float bPdf(const Bsdf& b, const BRec& r) {
    if constexpr (T == BTE_BSDF_DIFFUSE) return pdfDiffuse(r);
    else if constexpr (T == BTE_BSDF_MICROFACET) return pdfMicrofacet(b, r);
    else if constexpr (T == BTE_BSDF_DISNEY) return pdfDisney(b, r);
    else if constexpr (T == BTE_BSDF_DIELECTRIC || T == BTE_BSDF_MIRROR) return 0.0f;
    else return bsdfPdf(b, r);
}
// Sample and the pdf a separate bsdfPdf call on the sampled record returns
// (rt_path_integrator.cuh:37-39); where the sampler already evaluated that pdf
// on the final record the value is reused (same inputs, same result).
template <int T>
BTE_DEV V3 bSample(const Bsdf& b, BRec& r, float u1, float u2, float& pdf) {
    if constexpr (T == BTE_BSDF_DISNEY) {
        const float specT = b.p1.z;
        float ax, ay;
        disneyAlpha(b, ax, ay);
        if (r.wi.z <= 0.0f) {
            if (specT <= 0.0f) { pdf = 0.0f; return {0, 0, 0}; }
            sampleGlassLobe(b, r, u1, u2, ax, ay);
        } else {
            sampleDisneyLobeChoice(b, r, u1, u2, ax, ay);
        }
        pdf = pdfDisney(b, r);
        if (pdf < 1e-7f) return {0, 0, 0};
        return div3(evalDisney(b, r), pdf);
    } else if constexpr (T == BTE_BSDF_MICROFACET) {
        const V3 w = sampleMicrofacet(b, r, u1, u2);
        pdf = pdfMicrofacet(b, r);
        return w;
    } else {
        const V3 w = bsdfSample(b, r, u1, u2);
        pdf = bPdf<T>(b, r);
        return w;
    }
}

template <int T>
BTE_DEV NeeRec areaNEE(const DevScene& sc, V3 dIn, V3 thr, const Its& its, const Frame& f, const Bsdf& bsdf, float u0,
                       float u1) {
    NeeRec rec;
    rec.ok = false;
    if (sc.emitterCount <= 0) return rec;
    float epdf = 0.0f;
    const int eIdx = cdfReuse(sc.sceneCdf, sc.emitterCount, u0, epdf);
    const BteEmitter e = sc.emitters[eIdx];
    V3 pos = {0, 0, 0}, nrm = {0, 0, 0};
    float meshPdf = 0.0f;
    if (e.triCount > 0) {
        const float* triCdf = sc.emTriCdf + e.cdfOffset;
        float pick = 0.0f;
        const int li = cdfReuse(triCdf, e.triCount, u0, pick);
        const BteTriangle& tri = sc.tris[sc.emTriIdx[e.triIndexOffset + li]];
        const V3 v0 = {tri.v0.x, tri.v0.y, tri.v0.z}, v1 = {tri.v1.x, tri.v1.y, tri.v1.z}, v2 = {tri.v2.x, tri.v2.y, tri.v2.z};
        const float root = sqrtf(fmaxf(0.0f, 1.0f - u0));
        const float alpha = 1.0f - root;
        const float beta = u1 * root;
        const float gamma = 1.0f - alpha - beta;
        pos = add3(add3(mul3(v0, alpha), mul3(v1, beta)), mul3(v2, gamma));
        nrm = normalize3(cross3(sub3(v1, v0), sub3(v2, v0)));
        const float total = triCdf[e.triCount];
        if (total > 0.0f) meshPdf = 1.0f / total;
    }
    const V3 L = normalize3(sub3(pos, its.p));
    V3 le = {e.radiance.x, e.radiance.y, e.radiance.z};
    if (dot3(nrm, V3{-L.x, -L.y, -L.z}) <= 0.0f) le = {0, 0, 0};
    if ((le.x <= 0.0f && le.y <= 0.0f && le.z <= 0.0f) || epdf <= 0.0f || meshPdf <= 0.0f) return rec;
    const float dist = length3(sub3(pos, its.p));
    const float geo = fabsf(dot3(its.n, L)) * fabsf(dot3(nrm, mul3(L, -1.0f))) / (dist * dist);
    BRec br = makeRec();
    br.wi = toLocal(f, mul3(dIn, -1.0f));
    br.wo = toLocal(f, L);
    br.measure = 1;
    const V3 fr = bEval<T>(bsdf, br);
    const float lpa = epdf * meshPdf;
    const float lps = areaToSolidAngle(lpa, pos, its.p, nrm, L);
    const float bp = bPdf<T>(bsdf, br);
    const float den = bp + lps;
    if (den <= 0.0f) return rec;
    const float lw = lps / den;
    V3 c = mul3(fr, le);
    c = mul3(c, lw);
    c = mul3(c, geo);
    c = div3(c, lpa);
    rec.d = L;
    rec.tMax = dist - RT_EPS;
    rec.c = mul3(c, thr);
    rec.ok = true;
    return rec;
}

template <int T>
BTE_DEV NeeRec spotNEE(const DevScene& sc, V3 dIn, V3 thr, const Its& its, const Frame& f, const Bsdf& bsdf,
                       int si) {  // :175-230
    NeeRec rec;
    rec.ok = false;
    const BteSpotlight sp = sc.spots[si];
    const V3 pos = {sp.position.x, sp.position.y, sp.position.z};
    const V3 toL = sub3(pos, its.p);
    const float sq = dot3(toL, toL);
    if (sq < RT_EPS * RT_EPS) return rec;
    const float dist = sqrtf(sq);
    const V3 L = div3(toL, dist);
    const float lc = dot3(mul3(L, -1.0f), V3{sp.direction.x, sp.direction.y, sp.direction.z});
    if (lc <= sp.outerConeCosine) return rec;
    const V3 srad = {sp.radiance.x, sp.radiance.y, sp.radiance.z};
    V3 le;
    if (lc >= sp.innerConeCosine) {
        le = mul3(srad, sp.intensity / sq);
    } else {
        const float fo = (lc - sp.outerConeCosine) / (sp.innerConeCosine - sp.outerConeCosine);
        const float sm = (fo * fo) * (3.0f - (2.0f * fo));
        le = mul3(srad, (sp.intensity * sm) / sq);
    }
    BRec br = makeRec();
    br.wi = toLocal(f, mul3(dIn, -1.0f));
    br.wo = toLocal(f, L);
    br.measure = 1;
    const V3 fr = bEval<T>(bsdf, br);
    const float cs = fabsf(dot3(its.n, L));
    rec.d = L;
    rec.tMax = dist - RT_EPS;
    rec.c = mul3(mul3(mul3(fr, le), cs), thr);
    rec.ok = true;
    return rec;
}

template <int T>
BTE_DEV NeeRec envNEE(const DevScene& sc, V3 dIn, V3 thr, const Its& its, const Frame& f, const Bsdf& bsdf, float u0,
                      float u1) {
    NeeRec rec;
    rec.ok = false;
    if (sc.envTotal <= 0.0f) return rec;
    const int W = sc.envW, H = sc.envH;
    const float ms = sc.envMarg[H];
    if (ms <= 0.0f) return rec;
    const float xr = u0 * ms;
    const int row = cdfSearchGuided(sc.envMarg, H, xr, sc.envMargGuide, H, u0);
    const float c0r = sc.envMarg[row], c1r = sc.envMarg[row + 1];
    const float dRow = c1r - c0r;
    const float fRow = (dRow > 0.0f) ? (xr - c0r) / dRow : 0.0f;
    const float* rc = sc.envCond + (size_t)row * (size_t)(W + 1);
    const float rs = rc[W];
    if (rs <= 0.0f) return rec;
    const float xc = u1 * rs;
    const int col = cdfSearchGuided(rc, W, xc, sc.envCondGuide ? sc.envCondGuide + (size_t)row * (size_t)(W + 1) : nullptr,
                                    W, u1);
    const float c0c = rc[col], c1c = rc[col + 1];
    const float dCol = c1c - c0c;
    const float fCol = (dCol > 0.0f) ? (xc - c0c) / dCol : 0.0f;
    const float u = ((float)col + fCol) / (float)W;
    const float v = ((float)row + fRow) / (float)H;
    // equirectToDir (:56-67)
    const float phi = u * 2.0f * m::PI - m::PI - sc.envRot;
    const float theta = v * m::PI;
    float st, ct, sp, cp;
    dsincos(theta, &st, &ct);
    dsincos(phi, &sp, &cp);
    const V3 dir = normalize3({st * cp, ct, st * sp});
    const V3 tx = envBilinear(sc, u, v);
    const float k = sc.envIntensity;
    const V3 Le = {tx.x * k, tx.y * k, tx.z * k};
    const float pdfRow = dRow / ms;
    const float pdfCol = dCol / rs;
    const float s2 = fmaxf(dsin(theta), 1e-6f);
    const float pdf = (pdfRow * pdfCol * (float)W * (float)H) / (2.0f * m::PI * m::PI * s2);
    if (pdf <= 0.0f) return rec;
    if (Le.x <= 0.0f && Le.y <= 0.0f && Le.z <= 0.0f) return rec;
    const float ca = dot3(its.n, dir);
    if (ca <= 0.0f) return rec;
    BRec br = makeRec();
    br.wi = toLocal(f, mul3(dIn, -1.0f));
    br.wo = toLocal(f, dir);
    br.measure = 1;
    const V3 fr = bEval<T>(bsdf, br);
    const float bp = bPdf<T>(bsdf, br);
    const float sum = bp + pdf;
    if (sum <= 0.0f) return rec;
    const float lw = pdf / sum;
    V3 c = mul3(fr, Le);
    c = mul3(c, ca);
    c = div3(c, pdf);
    c = mul3(c, lw);
    rec.d = dir;
    rec.tMax = 1e30f - RT_EPS;
    rec.c = mul3(c, thr);
    rec.ok = true;
    return rec;
}

BTE_DEV void envHit(const DevScene& sc, V3 d, bool spec, float brdfPdf, V3 thr, V3& rad) {   // :237-259
    if (!sc.hasEnv) return;
    float u, v;
    dirToEquirect(d, sc.envRot, u, v);       // shared by eval and pdf (same inputs, same result)
    // one footprint record: the bilinear texels and the pdf factors of the
    // same four texels, one of which is the pdf's nearest texel
    const EnvFoot f = envFoot(sc, u, v);
    const float4* r = sc.envRec + 4 * f.rec;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    const V3 tx = envMix(f, r0, r1, r2);
    const V3 Le = {tx.x * sc.envIntensity, tx.y * sc.envIntensity, tx.z * sc.envIntensity};
    float w = 1.0f;
    if (!spec) {
        int ix, iy;
        envPdfTexel(sc, u, v, ix, iy);
        float q;
        if ((ix == f.ix0 || ix == f.ix1) && (iy == f.iy0 || iy == f.iy1)) {
            const float qy0 = (ix == f.ix0) ? r3.x : r3.y, qy1 = (ix == f.ix0) ? r3.z : r3.w;
            q = (iy == f.iy0) ? qy0 : qy1;
        } else {   // not reached for finite (u, v); the CDFs give the same value
            q = envPdfFactor(sc.envMarg, sc.envCond, sc.envW, sc.envH, sc.envTotal, ix, iy);
        }
        const float ep = envPdfFromFactor(q, v);
        const float sum = brdfPdf + ep;
        w = (sum > 0.0f) ? brdfPdf / sum : 0.0f;
    }
    rad = add3(rad, mul3(mul3(Le, thr), w));
}

// scatterPath (rt_path_integrator.cuh:25-47); writes the next ray, throughput,
// eta and brdfPdf.
template <int T>
BTE_DEV void scatter(const Its& its, const Frame& f, const Bsdf& bsdf, uint32_t& rng, V3& o, V3& d, V3& thr,
                     float& eta, float& brdfPdf) {
    BRec br = makeRec();
    br.wi = toLocal(f, mul3(d, -1.0f));
    br.rndExtra = rngNext(rng);
    const float u0 = rngNext(rng);
    const float u1 = rngNext(rng);
    const V3 w = bSample<T>(bsdf, br, u0, u1, brdfPdf);
    const V3 nd = normalize3(toWorld(f, br.wo));
    o = add3(its.p, mul3(nd, RT_EPS));
    d = nd;
    thr = mul3(thr, w);
    eta *= br.eta;
}
// russianRoulette (:54-65)
BTE_DEV bool roulette(int bounce, uint32_t& rng, V3& thr, float eta) {
    if (bounce > 3) {
        const float q = fminf(maxCoeff3(thr) * eta * eta, 0.99f);
        if (rngNext(rng) >= q) return false;
        thr = div3(thr, q);
    }
    return true;
}

BTE_DEV V3 primaryDir(const BteCamera& cam, int px, int py, int W, int H, float jx, float jy) {  // rt_camera.cuh:15-37
    const float fx = ((float)px + 0.5f + jx) / (float)W;
    const float fy = ((float)py + 0.5f + jy) / (float)H;
    const float ndcX = 2.0f * fx - 1.0f;
    const float ndcY = 2.0f * fy - 1.0f;
    const float th = dtan(0.5f * cam.fovYRadians);
    const float sx = ndcX * cam.aspect * th;
    const float sy = ndcY * th;
    const V3 fw = {cam.forward.x, cam.forward.y, cam.forward.z};
    const V3 rt = {cam.right.x, cam.right.y, cam.right.z};
    const V3 up = {cam.up.x, cam.up.y, cam.up.z};
    return normalize3(add3(fw, add3(mul3(rt, sx), mul3(up, sy))));
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, long long n, int v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Primary direction of a camera path and its random state after the pixel
// jitter draws (pixel / frame seed it).
BTE_DEV V3 cameraRay(const DevScene& sc, int pixel, int frame, uint32_t& rng) {
    const int px = pixel % sc.width, py = pixel / sc.width;
    rng = seedRng(hashUint((uint32_t)pixel ^ ((uint32_t)frame * 0x9e3779b9u)));
    const float jx = rngNext(rng) - 0.5f;
    const float jy = rngNext(rng) - 0.5f;
    return primaryDir(sc.cam, px, py, sc.width, sc.height, jx, jy);
}

// One camera path: queue slot `slot`, path id `pid`, pixel / frame seed it.
BTE_DEV void generateOne(const DevScene& sc, const WaveBuffers& wb, long long slot, uint32_t pid, int pixel,
                         int frame) {
    uint32_t rng;
    const V3 d = cameraRay(sc, pixel, frame, rng);
    wb.rayO[0][slot] = make_float4(sc.cam.origin.x, sc.cam.origin.y, sc.cam.origin.z, 0.0f);
    wb.rayD[0][slot] = make_float4(d.x, d.y, d.z, 1.0f);
    wb.aux[0][2 * slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(rng));
    wb.aux[0][2 * slot + 1] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(pid | kSpecBit));
    if (wb.aovAlb) {
        wb.aovAlb[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        wb.aovNrm[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// Queue order of a pass's paths: a wave's 64 slots hold Fg frames (16, 8, 4,
// 2 or 1: the largest that divides the pass's frame count) of a tile of
// 64 / Fg pixels (2x2 up to 8x8), so its primary rays start nearly alike;
// images whose sides are not multiples of the tile keep plain row order.
// Path ids stay frame * pixels + pixel; only the queue slots move.
BTE_DEV void slotPath(long long i, int pixels, int frames, int W, int H, int& frame, int& pixel) {
    const int fg = (frames % 16 == 0) ? 16 : (frames % 8 == 0) ? 8 : (frames % 4 == 0) ? 4 : (frames % 2 == 0) ? 2 : 1;
    const int tpx = 64 / fg;                                // pixels per tile: 8x8, 8x4, 4x4, 4x2, 2x2
    const int tw = (tpx >= 32) ? 8 : (tpx >= 8 ? 4 : 2), th = tpx / tw;
    if (W % tw || H % th) {
        frame = (int)(i / pixels);
        pixel = (int)(i % pixels);
        return;
    }
    const long long chunk = (long long)fg * pixels;
    const int c = (int)(i / chunk);
    const long long r = i - (long long)c * chunk;
    const long long group = r >> 6;
    const int within = (int)(r & 63);
    frame = c * fg + within % fg;
    const int inTile = within / fg;
    const int tilesX = W / tw;
    const int tx = (int)(group % tilesX), ty = (int)(group / tilesX);
    pixel = (ty * th + inTile / tw) * W + tx * tw + inTile % tw;
}

// slotPath for a pass below 2^31 slots (every single-scene pass: its counters
// are 32-bit), in 32-bit unsigned arithmetic: the same frame and pixel.
BTE_DEV void slotPath32(uint32_t i, uint32_t pixels, uint32_t frames, uint32_t W, uint32_t H, int& frame, int& pixel) {
    const uint32_t fg = (frames % 16u == 0) ? 16u : (frames % 8u == 0) ? 8u : (frames % 4u == 0) ? 4u : (frames % 2u == 0) ? 2u : 1u;
    const uint32_t tpx = 64u / fg;
    const uint32_t tw = (tpx >= 32u) ? 8u : (tpx >= 8u ? 4u : 2u), th = tpx / tw;
    if (W % tw || H % th) {
        frame = (int)(i / pixels);
        pixel = (int)(i % pixels);
        return;
    }
    const uint32_t chunk = fg * pixels;
    const uint32_t c = i / chunk;
    const uint32_t r = i - c * chunk;
    const uint32_t group = r >> 6, within = r & 63u;
    frame = (int)(c * fg + within % fg);
    const uint32_t inTile = within / fg;
    const uint32_t tilesX = W / tw;
    const uint32_t tx = group % tilesX, ty = group / tilesX;
    pixel = (int)((ty * th + inTile / tw) * W + tx * tw + inTile % tw);
}

__global__ __launch_bounds__(kBlock) void k_generate(DevScene sc, WaveBuffers wb, long long paths, int pixels) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= paths) return;
    const int frame0 = *wb.frameBase;
    int f, pixel;
    slotPath(i, pixels, (int)(paths / pixels), sc.width, sc.height, f, pixel);
    const uint32_t pid = (uint32_t)((long long)f * pixels + pixel);
    if (wb.camPaths == 0) {
        generateOne(sc, wb, i, pid, pixel, frame0 + f);
    } else {   // the state is computed where it is used; the rays only for per-ray traversal
        if (!wb.camRays) {
            uint32_t rng;
            const V3 d = cameraRay(sc, pixel, frame0 + f, rng);
            wb.rayO[0][i] = make_float4(sc.cam.origin.x, sc.cam.origin.y, sc.cam.origin.z, 0.0f);
            wb.rayD[0][i] = make_float4(d.x, d.y, d.z, 1.0f);
        }
        if (wb.aovAlb) {
            wb.aovAlb[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            wb.aovNrm[pid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    if (i == 0) wb.counters[kCtrRays] = (int)paths;
}

// Throughput, radiance and random state / path id of queue slot `slot` (its
// 32-byte record in wb.aux[cur]).
BTE_DEV void loadAux(const WaveBuffers& wb, int cur, long long slot, float4& th, float4& ra, uint2& st) {
    const float4 a0 = wb.aux[cur][2 * slot], a1 = wb.aux[cur][2 * slot + 1];
    ra = make_float4(a0.x, a0.y, a0.z, 0.0f);
    th = make_float4(a1.x, a1.y, a1.z, 0.0f);
    st = make_uint2(__float_as_uint(a0.w), __float_as_uint(a1.w));
}

// Bounce-0 state of queue slot `slot` when it is not stored (wb.camPaths >
// 0): the values generateOne writes for that slot.
BTE_DEV void cameraPath(const DevScene& sc, const WaveBuffers& wb, long long slot, float4& ro, float4& rd, float4& th,
                        float4& ra, uint2& st) {
    int f, pixel;
    slotPath32((uint32_t)slot, (uint32_t)wb.camPixels, (uint32_t)(wb.camPaths / wb.camPixels), (uint32_t)sc.width,
               (uint32_t)sc.height, f, pixel);
    uint32_t rng;
    const V3 d = cameraRay(sc, pixel, *wb.frameBase + f, rng);
    ro = make_float4(sc.cam.origin.x, sc.cam.origin.y, sc.cam.origin.z, 0.0f);
    rd = make_float4(d.x, d.y, d.z, 1.0f);
    th = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st = make_uint2(rng, (uint32_t)((long long)f * wb.camPixels + pixel) | kSpecBit);
}
// Path state of queue slot qi entering `bounce` (CAM: the bounce-0 camera path,
// computed).
template <bool CAM>
BTE_DEV void loadPath(const DevScene& sc, const WaveBuffers& wb, int bounce, int qi, float4& ro, float4& rd,
                      float4& th, float4& ra, uint2& st) {
    if constexpr (CAM) {
        cameraPath(sc, wb, qi, ro, rd, th, ra, st);
    } else {
        const int cur = bounce & 1;
        ro = wb.rayO[cur][qi];
        rd = wb.rayD[cur][qi];
        loadAux(wb, cur, qi, th, ra, st);
    }
}
// The ray of queue slot qi entering bounce 0 (the packet pass): computed when
// it is not stored.
BTE_DEV void loadCameraRay(const DevScene& sc, const WaveBuffers& wb, int qi, float4& ro, float4& rd) {
    if (wb.camPaths > 0 && wb.camRays) {
        int f, pixel;
        slotPath32((uint32_t)qi, (uint32_t)wb.camPixels, (uint32_t)(wb.camPaths / wb.camPixels), (uint32_t)sc.width,
                   (uint32_t)sc.height, f, pixel);
        uint32_t rng;
        const V3 d = cameraRay(sc, pixel, *wb.frameBase + f, rng);
        ro = make_float4(sc.cam.origin.x, sc.cam.origin.y, sc.cam.origin.z, 0.0f);
        rd = make_float4(d.x, d.y, d.z, 1.0f);
        return;
    }
    ro = wb.rayO[0][qi];
    rd = wb.rayD[0][qi];
}

// Batch: blockIdx.y = scene.  Scene s's paths take the path ids and queue
// slots of its range; pixel and frame seeds are the scene's own, so every
// scene renders exactly as it would alone.
__global__ __launch_bounds__(kBlock) void k_generate_batch(const DevScene* __restrict__ scenes, WaveBuffers wb,
                                                            long long totalSlots) {
    const int s = blockIdx.y;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const BatchSlot b = wb.slots[s];
    if (i >= b.paths) return;
    const DevScene sc = fetchScene(scenes, s);
    const int frame0 = *wb.frameBase;
    int f, pixel;
    slotPath(i, b.pixels, b.paths / b.pixels, sc.width, sc.height, f, pixel);
    generateOne(sc, wb, b.pathBase + i, (uint32_t)(b.pathBase + (long long)f * b.pixels + pixel), pixel,
                frame0 + f);
    if (i == 0) segCounts(wb, 0, kSegRays)[s] = b.paths;
    if (s == 0 && i == 0) wb.counters[kCtrRays] = (int)totalSlots;
}

// Category word of a finished extension ray: the shading category (0 = miss,
// else the BsdfType) in bits 0-2 and the hit triangle's BSDF index above, so
// the shading pass fetches the BSDF without going through the triangle.
BTE_DEV uint32_t shadeCategory(const DevScene& sc, int tri) {
    if (tri < 0) return kCatMiss;
    const int id = sc.triBsdf[tri];
    return ((uint32_t)id << 3) | (uint32_t)__float_as_int(sc.bsdfs[4 * (size_t)id + 3].w);
}

// Closest hit of every queued ray (src/wavefront/ray_extend.cu:49-72).  Waves
// take 64 consecutive queue slots per atomic: consecutive slots hold rays that
// were adjacent before compaction, so a wave's lanes share BVH cache lines.
template <bool COUNT, int L, int W>
BTE_DEV void extendOne(const DevScene& sc, const WaveBuffers& wb, int bounce, int qi, const StackRef<L>& stk,
                       uint32_t* nodes, uint32_t* tris) {
    const float4 o = wb.rayO[bounce & 1][qi];
    const float4 d = wb.rayD[bounce & 1][qi];
    float t = 0.0f, u = 0.0f, v = 0.0f;
    const int tri = traverseW<W, false, COUNT, L>(sc, xyz(o), xyz(d), RT_EPS, 1e30f, stk, t, u, v, nodes, tris);
    wb.hit[qi] = make_float4(__int_as_float(tri), t, u, v);
    wb.cat[qi] = shadeCategory(sc, tri);
}

template <bool COUNT, int L, int W, bool BATCH = false>
__global__ __launch_bounds__(kBlock) void k_extend(DevScene sc, WaveBuffers wb, int bounce,
                                                    const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : ctr[kCtrRays];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    uint32_t nodes = 0, tris = 0;
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkExtend);
    for (int chunk = first_chunk(grab); chunk < count; chunk = next_chunk(wk, grab, count)) {
        if constexpr (!BATCH) {
            for (int row = 0; row < wb.grabRows; ++row) {
                const int qi = chunk + row * 64 + lane_id();
                if (qi < count) extendOne<COUNT, L, W>(sc, wb, bounce, qi, stk, &nodes, &tris);
            }
        } else {
            const SegChunk k = segChunk(wb, bounce, kSegRays, chunk);
            if (k.left <= 0) continue;
            const DevScene S = fetchScene(scenes, k.s);
            for (int row = 0; row < wb.grabRows; ++row) {
                const int e = row * 64 + lane_id();
                if (e < k.left) extendOne<COUNT, L, W>(S, wb, bounce, k.first + e, stk, &nodes, &tris);
            }
        }
    }
    if (COUNT) {
        wave_add_u64(wb.travCtr + 0, nodes);
        wave_add_u64(wb.travCtr + 1, tris);
    }
}

// ---------------------------------------------------------------------------
// BVH4 / BVH8 passes with ray replacement.  Every lane of a persistent wave
// traces its own ray one node (or leaf triangle) per iteration; lanes whose ray has finished take
// the next slots of the wave's pool (its work chunks) once kRefill or more of
// them are idle, so the wave keeps its SIMD lanes busy instead of waiting for
// its slowest ray row by row.  In a batch the pool holds one scene at a time:
// a chunk of another scene is held until every lane has finished the current
// one, so the scene stays wave-uniform.  Results do not depend on the order in
// which rays or nodes are visited.
// ---------------------------------------------------------------------------
#ifndef BTE_REFILL
#define BTE_REFILL 24   // measured against 8, 16 and 32
#endif
constexpr int kRefill = BTE_REFILL;

// One closest-hit step: an inner node (hit children ordered by entry distance,
// the nearest next, the others pushed farthest first) or a leaf, then a pop.
// Returns true when the stack is empty: the query is done.
// Entry distances of the eight children of a BVH8 node (kNone: missed).
BTE_DEV void enter8(const DevScene& S, const WtRay& r, int node, float best, float (&k)[8], int (&c)[8]) {
    constexpr float kNone = __builtin_huge_valf();
    const float4* np = S.nodes8 + 16 * (size_t)node;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float4 lx = np[h], hx = np[2 + h], ly = np[4 + h], hy = np[6 + h], lz = np[8 + h], hz = np[10 + h];
        const int4 ch = *reinterpret_cast<const int4*>(np + 12 + h);
        float f0, f1, f2, f3;
        const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, RT_EPS, best, f0);
        const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, RT_EPS, best, f1);
        const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, RT_EPS, best, f2);
        const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, RT_EPS, best, f3);
        k[4 * h + 0] = e0 <= f0 ? e0 : kNone;
        k[4 * h + 1] = e1 <= f1 ? e1 : kNone;
        k[4 * h + 2] = e2 <= f2 ? e2 : kNone;
        k[4 * h + 3] = e3 <= f3 ? e3 : kNone;
        c[4 * h + 0] = ch.x; c[4 * h + 1] = ch.y; c[4 * h + 2] = ch.z; c[4 * h + 3] = ch.w;
    }
}

template <bool COUNT, int L, int W = 4>
BTE_DEV bool closestStep(const DevScene& S, const WtRay& r, int& node, int& sp, const StackRef<L>& stk, float& best,
                         int& bestTri, float& tO, float& uO, float& vO, uint32_t& nodes, uint32_t& tris) {
    constexpr float kNone = __builtin_huge_valf();
    if (W == 8 && node >= 0) {
        if (COUNT) ++nodes;
        float k[8];
        int c[8];
        enter8(S, r, node, best, k, c);
        // 19-comparator sorting network (nearest first)
        sortPair(k[0], c[0], k[2], c[2]); sortPair(k[1], c[1], k[3], c[3]);
        sortPair(k[4], c[4], k[6], c[6]); sortPair(k[5], c[5], k[7], c[7]);
        sortPair(k[0], c[0], k[4], c[4]); sortPair(k[1], c[1], k[5], c[5]);
        sortPair(k[2], c[2], k[6], c[6]); sortPair(k[3], c[3], k[7], c[7]);
        sortPair(k[0], c[0], k[1], c[1]); sortPair(k[2], c[2], k[3], c[3]);
        sortPair(k[4], c[4], k[5], c[5]); sortPair(k[6], c[6], k[7], c[7]);
        sortPair(k[2], c[2], k[4], c[4]); sortPair(k[3], c[3], k[5], c[5]);
        sortPair(k[1], c[1], k[4], c[4]); sortPair(k[3], c[3], k[6], c[6]);
        sortPair(k[1], c[1], k[2], c[2]); sortPair(k[3], c[3], k[4], c[4]); sortPair(k[5], c[5], k[6], c[6]);
        if (k[0] != kNone) {
#pragma unroll
            for (int j = 7; j >= 1; --j)
                if (k[j] != kNone) push(stk, sp, c[j]);
            node = c[0];
            return false;
        }
    } else if (node >= 0) {
        if (COUNT) ++nodes;
        const float4* np = S.nodes4 + 8 * (size_t)node;
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const int4 ch = *reinterpret_cast<const int4*>(np + 6);
        float f0, f1, f2, f3;
        const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, RT_EPS, best, f0);
        const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, RT_EPS, best, f1);
        const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, RT_EPS, best, f2);
        const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, RT_EPS, best, f3);
        float k0 = e0 <= f0 ? e0 : kNone, k1 = e1 <= f1 ? e1 : kNone;
        float k2 = e2 <= f2 ? e2 : kNone, k3 = e3 <= f3 ? e3 : kNone;
        int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
        sortPair(k0, c0, k1, c1);
        sortPair(k2, c2, k3, c3);
        sortPair(k0, c0, k2, c2);
        sortPair(k1, c1, k3, c3);
        sortPair(k1, c1, k2, c2);
        if (k0 != kNone) {
            push3(stk, sp, c1, k1 != kNone, c2, k2 != kNone, c3, k3 != kNone);
            node = c0;
            return false;
        }
    } else {
        const int code = ~node;
        const int start = code >> kLeafCountBits;
        const int cnt = (code & (kMaxLeafTris - 1)) + 1;
        // one triangle per step: an iteration of the wave costs one triangle
        // test for its lanes at leaves, not the longest leaf's loop
        {
            const float4* tp = S.leafTris + 3 * (size_t)start;
            const float4 a = tp[0], b = tp[1], c = tp[2];
            if (COUNT) ++tris;
            float t, u, v;
            if (hitTriangle(r, a, b, c, RT_EPS, best, t, u, v)) {
                const int id = __float_as_int(a.w);
                if (t < best || bestTri < 0 || id < bestTri) {
                    best = t; bestTri = id; tO = t; uO = u; vO = v;
                }
            }
            if (cnt > 1) {   // the rest of the leaf is the next step
                node = ~(((start + 1) << kLeafCountBits) | (cnt - 2));
                return false;
            }
        }
    }
    if (sp == 0) return true;
    node = pop(stk, sp);
    return false;
}

// One any-hit step (children in any order, a hit child first); returns 1 when
// a triangle blocks the ray, 2 when the stack is empty (unblocked), else 0.
template <bool COUNT, int L, int W = 4>
BTE_DEV int anyStep(const DevScene& S, const WtRay& r, float tmax, int& node, int& sp, const StackRef<L>& stk,
                    uint32_t& nodes, uint32_t& tris) {
    constexpr float kNone = __builtin_huge_valf();
    if (W == 8 && node >= 0) {
        if (COUNT) ++nodes;
        float k[8];
        int c[8];
        enter8(S, r, node, tmax, k, c);
        // any hit ends the search: descend into the first hit child, push the others
        int first = -1;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (first < 0 && k[j] != kNone) first = j;
        if (first >= 0) {
#pragma unroll
            for (int j = 7; j >= 0; --j)
                if (j != first && k[j] != kNone) push(stk, sp, c[j]);
            node = c[first];
            return 0;
        }
    } else if (node >= 0) {
        if (COUNT) ++nodes;
        const float4* np = S.nodes4 + 8 * (size_t)node;
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const int4 ch = *reinterpret_cast<const int4*>(np + 6);
        float f0, f1, f2, f3;
        const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, RT_EPS, tmax, f0);
        const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, RT_EPS, tmax, f1);
        const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, RT_EPS, tmax, f2);
        const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, RT_EPS, tmax, f3);
        float k0 = e0 <= f0 ? e0 : kNone, k1 = e1 <= f1 ? e1 : kNone;
        float k2 = e2 <= f2 ? e2 : kNone, k3 = e3 <= f3 ? e3 : kNone;
        int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
        if (k0 == kNone) {
            if (k1 != kNone) { k0 = k1; c0 = c1; k1 = kNone; }
            else if (k2 != kNone) { k0 = k2; c0 = c2; k2 = kNone; }
            else if (k3 != kNone) { k0 = k3; c0 = c3; k3 = kNone; }
        }
        if (k0 != kNone) {
            push3(stk, sp, c1, k1 != kNone, c2, k2 != kNone, c3, k3 != kNone);
            node = c0;
            return 0;
        }
    } else {
        const int code = ~node;
        const int start = code >> kLeafCountBits;
        const int cnt = (code & (kMaxLeafTris - 1)) + 1;
        // one triangle per step: an iteration of the wave costs one triangle
        // test for its lanes at leaves, not the longest leaf's loop
        {
            const float4* tp = S.leafTris + 3 * (size_t)start;
            const float4 a = tp[0], b = tp[1], c = tp[2];
            if (COUNT) ++tris;
            float t, u, v;
            if (hitTriangle(r, a, b, c, RT_EPS, tmax, t, u, v)) return 1;
            if (cnt > 1) {   // the rest of the leaf is the next step
                node = ~(((start + 1) << kLeafCountBits) | (cnt - 2));
                return 0;
            }
        }
    }
    if (sp == 0) return 2;
    node = pop(stk, sp);
    return 0;
}

// The wave's pool of queue slots: [beg, end) of the current scene, plus one
// grabbed chunk [nBeg, nEnd) of scene nS not yet started.  All fields are
// wave-uniform.
struct RayPool {
    int beg, end, nBeg, nEnd, nS, curS, chunk;
    bool held, exhausted;
};

BTE_DEV RayPool makePool(int grab) {
    RayPool p;
    p.beg = p.end = p.nBeg = p.nEnd = p.nS = 0;
    p.curS = -1;
    p.chunk = first_chunk(grab);
    p.held = p.exhausted = false;
    return p;
}

// Hands the idle lanes the pool's next slots: start(slot) prepares a lane's
// query and returns whether it needs traversal (false: finished at once).
// Returns the idle mask afterwards; all idle = nothing left to start.
template <bool BATCH, class Start>
BTE_DEV unsigned long long refill(RayPool& p, DevScene& S, const DevScene* __restrict__ scenes, const WaveBuffers& wb,
                                  int bounce, int q, int* wk, int grab, int count, bool& active, Start start) {
    unsigned long long idle = __ballot(!active);
    while (idle != 0ull) {
        if (p.beg >= p.end) {
            if (!p.held) {
                if (p.exhausted) break;
                const int c = p.chunk >= 0 ? p.chunk : next_chunk(wk, grab, count);
                p.chunk = -1;
                if (c >= count) {
                    p.exhausted = true;
                    break;
                }
                if constexpr (BATCH) {
                    const SegChunk k = segChunk(wb, bounce, q, c);
                    if (k.left <= 0) continue;
                    p.nBeg = k.first;
                    p.nEnd = k.first + min(grab, k.left);
                    p.nS = k.s;
                } else {
                    p.nBeg = c;
                    p.nEnd = min(c + grab, count);
                }
                p.held = true;
            }
            if constexpr (BATCH) {
                if (p.nS != p.curS) {
                    if (idle != ~0ull) break;    // lanes still trace the current scene
                    p.curS = p.nS;
                    S = fetchScene(scenes, p.curS);
                }
            }
            p.beg = p.nBeg;
            p.end = p.nEnd;
            p.held = false;
        }
        const unsigned lo = (unsigned)(idle & 0xffffffffull), hi = (unsigned)(idle >> 32);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        const int n = __popcll(idle);
        if (!active && rank < p.end - p.beg) active = start(p.beg + rank);
        p.beg = min(p.beg + n, p.end);
        idle = __ballot(!active);
    }
    return idle;
}

template <bool COUNT, int L, bool BATCH, int W = 4>
__global__ __launch_bounds__(kBlock) void k_extend_pool(DevScene sc, WaveBuffers wb, int bounce,
                                                         const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : ctr[kCtrRays];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    uint32_t nodes = 0, tris = 0;
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkExtend);
    RayPool pool = makePool(grab);
    DevScene S = sc;
    bool active = false;
    int qi = 0, node = 0, sp = 0, bestTri = -1;
    float best = 0.0f, tO = 0.0f, uO = 0.0f, vO = 0.0f;
    WtRay r;
    auto start = [&](int slot) {
        qi = slot;
        const float4 o = wb.rayO[bounce & 1][qi], d = wb.rayD[bounce & 1][qi];
        r = makeWtRay(xyz(o), xyz(d));
        best = 1e30f;
        bestTri = -1;
        tO = uO = vO = 0.0f;
        node = 0;
        sp = 0;
        if (r.valid && S.triCount > 0) return true;
        wb.hit[qi] = make_float4(__int_as_float(-1), 0.0f, 0.0f, 0.0f);   // nothing to traverse: a miss
        wb.cat[qi] = shadeCategory(S, -1);
        return false;
    };
    for (;;) {
        const int nIdle = __popcll(__ballot(!active));
        if (nIdle >= kRefill || nIdle == 64) {
            if (refill<BATCH>(pool, S, scenes, wb, bounce, kSegRays, wk, grab, count, active, start) == ~0ull) break;
        }
        if (!active) continue;
        if (closestStep<COUNT, L, W>(S, r, node, sp, stk, best, bestTri, tO, uO, vO, nodes, tris)) {
            wb.hit[qi] = make_float4(__int_as_float(bestTri), tO, uO, vO);
            wb.cat[qi] = shadeCategory(S, bestTri);
            active = false;
        }
    }
    if (COUNT) {
        wave_add_u64(wb.travCtr + 0, nodes);
        wave_add_u64(wb.travCtr + 1, tris);
    }
}

// Camera rays as packets.  A wave's 64 queue slots on the first bounce are 16
// frames of one 2x2 pixel tile: nearly the same ray.  The wave traverses the
// BVH4 once for all of them: a child is entered when any lane's ray enters it
// before that lane's closest hit so far, children go in order of the entry
// distance of their first entering lane, the stack is the wave's own (LDS,
// uniform) and node and triangle loads are wave-uniform (scalar).  Every node
// that a lane's own traversal would have to visit to find its closest hit is
// visited, and each lane runs its own triangle tests, so the hits are those
// of the per-ray traversal (the closest hit does not depend on the set or
// order of the visited nodes).  No lane diverges from the wave's step.
template <class T>
BTE_DEV const T* constPtr(const T* p) {
    return (const T*)(const __attribute__((address_space(4))) T*)(uintptr_t)p;
}
BTE_DEV float uniformLane(float v, unsigned long long mask) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)__ffsll((long long)mask) - 1));
}

// In a batch a chunk (one compact row) lies in one scene.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_extend_packet(DevScene sc, WaveBuffers wb, int bounce,
                                                           const DevScene* __restrict__ scenes) {
    constexpr float kNone = __builtin_huge_valf();
    __shared__ int stacks[kBlock / 64][kStackSize];
    int* stk = stacks[threadIdx.x >> 6];
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : wb.counters[(size_t)bounce * kCtrStride + kCtrRays];
    int* wk = work_group(wb, bounce, kWorkExtend);
    for (int chunk = first_chunk(64); chunk < count; chunk = next_chunk(wk, 64, count)) {
        int first = chunk, left = count - chunk;
        if constexpr (BATCH) {
            const SegChunk k = segChunk(wb, bounce, kSegRays, chunk);
            if (k.left <= 0) continue;
            first = k.first;
            left = k.left;
            sc = fetchScene(scenes, k.s);
        }
        const float4* nodes = constPtr(sc.nodes4);
        const float4* tris = constPtr(sc.leafTris);
        const int qi = first + lane_id();
        const bool valid = lane_id() < left;
        V3 o = {0, 0, 0}, d = {0, 0, 1};
        if (valid) {
            float4 ro, rd;
            loadCameraRay(sc, wb, qi, ro, rd);
            o = xyz(ro);
            d = xyz(rd);
        }
        const WtRay r = makeWtRay(o, d);
        const bool act = valid && r.valid && sc.triCount > 0;
        float best = 1e30f, tO = 0.0f, uO = 0.0f, vO = 0.0f;
        int bestTri = -1;
        int node = 0, sp = 0;
        bool more = __ballot(act) != 0ull;
        while (more) {
            if (node >= 0) {
                const float4* np = nodes + 8 * (size_t)node;
                const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
                const int4 ch = *reinterpret_cast<const int4*>(np + 6);
                float f0, f1, f2, f3;
                const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, RT_EPS, best, f0);
                const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, RT_EPS, best, f1);
                const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, RT_EPS, best, f2);
                const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, RT_EPS, best, f3);
                const unsigned long long m0 = __ballot(act && e0 <= f0), m1 = __ballot(act && e1 <= f1);
                const unsigned long long m2 = __ballot(act && e2 <= f2), m3 = __ballot(act && e3 <= f3);
                float k0 = m0 ? uniformLane(e0, m0) : kNone, k1 = m1 ? uniformLane(e1, m1) : kNone;
                float k2 = m2 ? uniformLane(e2, m2) : kNone, k3 = m3 ? uniformLane(e3, m3) : kNone;
                int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
                sortPair(k0, c0, k1, c1);
                sortPair(k2, c2, k3, c3);
                sortPair(k0, c0, k2, c2);
                sortPair(k1, c1, k3, c3);
                sortPair(k1, c1, k2, c2);
                if (k0 != kNone) {
                    // farthest first, so the nearer children are popped first
                    if (k3 != kNone) stk[sp++] = c3;
                    if (k2 != kNone) stk[sp++] = c2;
                    if (k1 != kNone) stk[sp++] = c1;
                    node = c0;
                    continue;
                }
            } else {
                const int code = ~node;
                const int start = code >> kLeafCountBits;
                const int cnt = (code & (kMaxLeafTris - 1)) + 1;
                for (int j = 0; j < cnt; ++j) {
                    const float4* tp = tris + 3 * (size_t)(start + j);
                    const float4 a = tp[0], b = tp[1], c = tp[2];
                    float t, u, v;
                    if (act && hitTriangle(r, a, b, c, RT_EPS, best, t, u, v)) {
                        const int id = __float_as_int(a.w);
                        if (t < best || bestTri < 0 || id < bestTri) {
                            best = t; bestTri = id; tO = t; uO = u; vO = v;
                        }
                    }
                }
            }
            if (sp == 0) break;
            node = __builtin_amdgcn_readfirstlane(stk[--sp]);
        }
        if (valid) {
            wb.hit[qi] = make_float4(__int_as_float(bestTri), tO, uO, vO);
            wb.cat[qi] = shadeCategory(sc, bestTri);
        }
    }
}

// Groups the queue slots of one bounce by shading category, keeping queue
// order within each category.  A block handles 1024 consecutive slots per
// iteration: ballot ranks within each wave, an LDS scan over the block, and one
// global atomic per non-empty category (device-wide atomics on a hot counter
// serialise, so they are amortised over the block).
// A missed ray ends its path: env-map emission with MIS (A13) added to its
// radiance, which becomes the path's final radiance.
template <bool CAM>
BTE_DEV void shadeMiss(const DevScene& sc, const WaveBuffers& wb, int qi, int bounce) {
    float4 ro, rd, th, ra;
    uint2 st;
    loadPath<CAM>(sc, wb, bounce, qi, ro, rd, th, ra, st);
    const uint32_t pid = st.y & kPidMask;
    const bool spec = (st.y & kSpecBit) != 0;
    if (wb.hitCapture && bounce < wb.maxRec)
        wb.hitCapture[(size_t)pid * wb.maxRec + bounce] = __float_as_int(wb.hit[qi].x);
    V3 rad = xyz(ra);
    envHit(sc, xyz(rd), spec, ro.w, xyz(th), rad);
    wb.finalRad[pid] = make_float4(rad.x, rad.y, rad.z, 0.0f);
}

// 64-slot rows per wave per iteration: 8 for a single scene; 4 in a batch,
// whose per-scene queue segments are aligned so that a block's 4 x 256
// entries lie in one scene
template <bool BATCH>
constexpr int catSub() { return BATCH ? 4 : 8; }
static_assert(catSub<true>() * kBlock <= kSegRowAlign * 64, "a batch categoriser block must lie in one scene segment");
// Misses are shaded here, in queue order, and are not queued.
template <bool BATCH = false, bool CAM = false>
__global__ __launch_bounds__(kBlock) void k_categorize(DevScene sc, WaveBuffers wb, int bounce,
                                                        const DevScene* __restrict__ scenes) {
    constexpr int kCatSub = catSub<BATCH>();
    __shared__ int sCnt[kCatSub][kBlock / 64][kNumCats];
    __shared__ int sBase[kNumCats];
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : ctr[kCtrRays];
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const int per = kCatSub * kBlock;
    for (int base = blockIdx.x * per; base < count; base += gridDim.x * per) {
        // batch: a 1024-entry block of compact rows lies in one scene
        int first = base, left = count - base, begin = 0, s = 0;
        if constexpr (BATCH) {
            const SegChunk k = segChunk(wb, bounce, kSegRays, base);
            if (k.left <= 0) continue;
            s = k.s;
            first = k.first;
            left = k.left;
            begin = (int)wb.slots[s].pathBase;
        }
        int c[kCatSub];
        for (int j = 0; j < kCatSub; ++j) {
            const int e = j * kBlock + wave * 64 + lane;
            const int i = first + e;
            c[j] = (e < left) ? (int)(wb.cat[i] & 7u) : -1;
            if (c[j] == kCatMiss) {
                if constexpr (BATCH) shadeMiss<false>(fetchScene(scenes, s), wb, i, bounce);
                else shadeMiss<CAM>(sc, wb, i, bounce);
                c[j] = -1;
            }
            for (int k = 1; k < kNumCats; ++k) {
                const unsigned long long m = __ballot(c[j] == k);
                if (lane == 0) sCnt[j][wave][k] = __popcll(m);
            }
        }
        __syncthreads();
        if (threadIdx.x >= 1 && threadIdx.x < kNumCats) {
            const int k = threadIdx.x;
            int run = 0;
            for (int j = 0; j < kCatSub; ++j)
                for (int w = 0; w < kBlock / 64; ++w) {
                    const int n = sCnt[j][w][k];
                    sCnt[j][w][k] = run;
                    run += n;
                }
            int* catCtr = BATCH ? segCounts(wb, bounce, kSegCat + k) + s : ctr + kCtrCat + k;
            sBase[k] = run ? begin + atomicAdd(catCtr, run) : 0;
            // a single scene's hits are the sum of its category counts (added
            // up on the host): no second atomic on a shared counter
            if (BATCH && run) atomicAdd(ctr + kCtrHits, run);
        }
        __syncthreads();
        for (int j = 0; j < kCatSub; ++j) {
            const int i = first + j * kBlock + wave * 64 + lane;
            for (int k = 1; k < kNumCats; ++k) {
                const unsigned long long m = __ballot(c[j] == k);
                if (c[j] == k) {
                    const unsigned lo = (unsigned)(m & 0xffffffffull), hi = (unsigned)(m >> 32);
                    const int rank = (int)__builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
                    wb.catQ[(size_t)k * (size_t)wb.cap + (size_t)(sBase[k] + sCnt[j][wave][k] + rank)] = i;
                }
            }
        }
        __syncthreads();
    }
}

// Where a shading row appends: surviving paths at survBase + (*survCtr)++,
// shadow records of light class k at shBase + (*(shCtr + k * shStride))++.
struct ShadeSinks {
    int* survCtr;
    int survBase;
    int* shCtr;
    int shStride;
    int shBase;
};

// Shades one 64-slot row of a category queue: lanes with `active` set hold
// the hits at list[li].  Every lane of the block must call it the same number
// of times (block-wide appends).
template <int T, bool BATCH, bool CAM>
BTE_DEV void shadeRow(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, int li, bool active,
                      const ShadeSinks& out, BlockAppendLds& blk) {
    const int cur = bounce & 1, nxt = cur ^ 1;
    const int* list = wb.catQ + (size_t)T * (size_t)wb.cap;
    constexpr bool kDiffuseType = (T == BTE_BSDF_DIFFUSE || T == BTE_BSDF_MICROFACET || T == BTE_BSDF_DISNEY);
    V3 o = {0, 0, 0}, d = {0, 0, 0}, thr = {0, 0, 0}, rad = {0, 0, 0};
    float brdfPdf = 0.0f, eta = 0.0f;
    uint32_t rng = 0, pid = 0, aovBit = 0;
    bool spec = false;
    float4 h = make_float4(0, 0, 0, 0);
    int bsdfId = 0;
    if (active) {
        const int qi = list[li];
        float4 ro, rd, th, ra;
        uint2 st;
        loadPath<CAM>(sc, wb, bounce, qi, ro, rd, th, ra, st);
        h = wb.hit[qi];
        bsdfId = (int)(wb.cat[qi] >> 3);
        o = xyz(ro); brdfPdf = ro.w;
        d = xyz(rd); eta = rd.w;
        thr = xyz(th);
        rad = xyz(ra);
        rng = st.x;
        pid = st.y & kPidMask;
        aovBit = st.y & kAovBit;
        spec = (st.y & kSpecBit) != 0;
        if (wb.hitCapture && bounce < wb.maxRec) wb.hitCapture[(size_t)pid * wb.maxRec + bounce] = __float_as_int(h.x);
    }
    bool survive = false;
    Its its;
    its.tri = -1; its.p = {0, 0, 0}; its.n = {0, 0, 0};
    Bsdf bsdf;
    bsdf.type = T;
    Frame frame;
    frame.t = frame.b = frame.n = V3{0, 0, 0};
    const V3 dIn = d, thrIn = thr;
    const uint32_t rngNee = rng;
    if (active) {
        its = makeIts(sc, o, d, h);
        bsdf = loadBsdfId(sc, bsdfId);
        applyTextures(sc, its.tri, h.z, h.w, bsdf);
        if (sc.triFlags && sc.triFlags[its.tri] != 0) emitterHit(sc, o, d, spec, brdfPdf, thr, its, rad);
        spec = !kDiffuseType;
        if (kDiffuseType) {
            rngNext(rng); rngNext(rng);                       // area NEE draws (always)
            if (sc.hasEnv) { rngNext(rng); rngNext(rng); }    // env NEE draws
        }
        if (kDiffuseType && wb.aovAlb && !aovBit) {
            // first non-delta hit: base colour and shading normal
            wb.aovAlb[pid] = make_float4(bsdf.p0.x, bsdf.p0.y, bsdf.p0.z, 0.0f);
            wb.aovNrm[pid] = make_float4(its.n.x, its.n.y, its.n.z, 0.0f);
            aovBit = kAovBit;
        }
        frame = makeFrame(its.n);
        scatter<T>(its, frame, bsdf, rng, o, d, thr, eta, brdfPdf);
        survive = roulette(bounce, rng, thr, eta) && (bounce + 1 < maxDepth);
    }
    const int slot = out.survBase + block_append<BATCH>(out.survCtr, survive, blk.n, blk.ctrs);
    if (survive) {
        wb.rayO[nxt][slot] = make_float4(o.x, o.y, o.z, brdfPdf);
        wb.rayD[nxt][slot] = make_float4(d.x, d.y, d.z, eta);
        wb.aux[nxt][2 * (size_t)slot] = make_float4(rad.x, rad.y, rad.z, __uint_as_float(rng));
        wb.aux[nxt][2 * (size_t)slot + 1] =
            make_float4(thr.x, thr.y, thr.z, __uint_as_float(pid | aovBit | (spec ? kSpecBit : 0u)));
    } else if (active) {
        wb.finalRad[pid] = make_float4(rad.x, rad.y, rad.z, 0.0f);
    }
    if constexpr (kDiffuseType) {
        // NEE records, one light class at a time in the reference's order
        const uint32_t dest = survive ? (uint32_t)slot : (pid | kFinalBit);
        uint32_t r2 = rngNee;
        float a0 = 0.0f, a1 = 0.0f, e0 = 0.0f, e1 = 0.0f;
        if (active) {
            a0 = rngNext(r2); a1 = rngNext(r2);
            if (sc.hasEnv) { e0 = rngNext(r2); e1 = rngNext(r2); }
        }
        for (int k = 0; k < wb.kinds; ++k) {
            NeeRec rec;
            rec.ok = false;
            const bool spots = (k == wb.kindSpot);
            if (active) {
                // a batch's light classes are the union over its scenes
                if (k == wb.kindArea) rec = areaNEE<T>(sc, dIn, thrIn, its, frame, bsdf, a0, a1);
                else if (k == wb.kindEnv) { if (sc.hasEnv) rec = envNEE<T>(sc, dIn, thrIn, its, frame, bsdf, e0, e1); }
                else if (spots) rec.ok = sc.spotCount > 0;     // a request: the spot pass evaluates every spotlight
            }
            const int sl = out.shBase + block_append<BATCH>(out.shCtr + k * out.shStride, rec.ok, blk.n, blk.ctrs);
            if (rec.ok) {
                const size_t idx = (size_t)k * (size_t)wb.cap + (size_t)sl;
                if (spots) {
                    wb.shO[idx] = make_float4(its.p.x, its.p.y, its.p.z, __uint_as_float(dest));
                    wb.shD[idx] = make_float4(its.n.x, its.n.y, its.n.z, __int_as_float(its.tri));
                    wb.shC[idx] = make_float4(dIn.x, dIn.y, dIn.z, h.z);
                    wb.shX[sl] = make_float4(thrIn.x, thrIn.y, thrIn.z, h.w);
                } else {
                    wb.shO[idx] = make_float4(its.p.x, its.p.y, its.p.z, rec.tMax);
                    wb.shD[idx] = make_float4(rec.d.x, rec.d.y, rec.d.z, __uint_as_float(dest));
                    wb.shC[idx] = make_float4(rec.c.x, rec.c.y, rec.c.z, 0.0f);
                }
            }
        }
    }
}

// Shade one category of hits (T = BsdfType; misses are shaded by k_categorize).
// Registers held to 4 waves per SIMD (128): the rows wait on chains of
// gathers, and the Disney kernel (142 registers unforced, 3 waves) runs
// faster with the extra wave than without its few spills.
template <int T, bool BATCH = false, bool CAM = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(DevScene sc, WaveBuffers wb,
                                                                                          int bounce, int maxDepth,
                                                   const DevScene* __restrict__ scenes) {
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    int* ctrNext = wb.counters + (size_t)(bounce + 1) * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegCat + T) : ctr[kCtrCat + T];
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkShade + T);
    // the block's waves shade their rows in step and append block-wide: one
    // atomic per block, row and counter (a batch's counters are per scene;
    // waves on the same scene share one)
    __shared__ BlockAppendLds blk;
    for (int chunk = first_chunk(grab);;) {
        const bool has = chunk < count;
        if (!__syncthreads_or(has)) break;
        if constexpr (!BATCH) {
            const ShadeSinks out{ctrNext + kCtrRays, 0, ctr + kCtrShadow, 1, 0};
            for (int row = 0; row < wb.grabRows; ++row) {
                const int li = chunk + row * 64 + lane_id();
                shadeRow<T, false, CAM>(sc, wb, bounce, maxDepth, li, has && li < count, out, blk);
            }
        } else {
            SegChunk k{0, 0, 0};   // a wave without a chunk shades nothing (scene 0 stands in)
            if (has) k = segChunk(wb, bounce, kSegCat + T, chunk);
            // wave-uniform (the scene descriptor stays in scalar registers)
            const int sIdx = __builtin_amdgcn_readfirstlane(k.s);
            const int begin = (int)wb.slots[sIdx].pathBase;
            const DevScene S = fetchScene(scenes, sIdx);
            const ShadeSinks out{segCounts(wb, bounce + 1, kSegRays) + sIdx, begin,
                                 segCounts(wb, bounce, kSegShadow) + sIdx, wb.sceneCount, begin};
            for (int row = 0; row < wb.grabRows; ++row) {
                const int e = row * 64 + lane_id();
                shadeRow<T, true, false>(S, wb, bounce, maxDepth, k.first + e, has && e < k.left, out, blk);
            }
        }
        if (has) chunk = next_chunk(wk, grab, count);
    }
}

template <bool COUNT, int L, int W>
BTE_DEV void connectOne(const DevScene& sc, const WaveBuffers& wb, int nxt, size_t si, const StackRef<L>& stk,
                        uint32_t* nodes, uint32_t* tris) {
    const float4 so = wb.shO[si];
    const float4 sd = wb.shD[si];
    float t, u, v;
    const int occ = traverseW<W, true, COUNT, L>(sc, xyz(so), xyz(sd), RT_EPS, so.w, stk, t, u, v, nodes, tris);
    if (!occ) {
        const float4 c = wb.shC[si];
        const uint32_t dest = __float_as_uint(sd.w);
        float4* r = (dest & kFinalBit) ? (wb.finalRad + (dest & ~kFinalBit)) : (wb.aux[nxt] + 2 * (size_t)dest);
        float4 v4 = *r;
        v4.x = v4.x + c.x;
        v4.y = v4.y + c.y;
        v4.z = v4.z + c.z;
        *r = v4;
    }
}

// Shadow-ray (any-hit) pass of one light class with ray replacement.  An
// unoccluded ray adds its contribution to its path (each path has at most one
// shadow ray per light class, so no two lanes write the same path).
template <bool COUNT, int L, bool BATCH, int W = 4>
__global__ __launch_bounds__(kBlock) void k_connect_pool(DevScene sc, WaveBuffers wb, int bounce, int kind,
                                                          const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    const int nxt = (bounce & 1) ^ 1;
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegShadow + kind) : ctr[kCtrShadow + kind];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    const size_t off = (size_t)kind * (size_t)wb.cap;
    uint32_t nodes = 0, tris = 0;
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkConnect + kind);
    RayPool pool = makePool(grab);
    DevScene S = sc;
    bool active = false;
    size_t si = 0;
    int node = 0, sp = 0;
    float tmax = 0.0f;
    WtRay r;
    auto unoccluded = [&](size_t slot) {
        const float4 c = wb.shC[slot];
        const uint32_t dest = __float_as_uint(wb.shD[slot].w);
        float4* rp = (dest & kFinalBit) ? (wb.finalRad + (dest & ~kFinalBit)) : (wb.aux[nxt] + 2 * (size_t)dest);
        float4 v4 = *rp;
        v4.x = v4.x + c.x;
        v4.y = v4.y + c.y;
        v4.z = v4.z + c.z;
        *rp = v4;
    };
    auto start = [&](int slot) {
        si = off + (size_t)slot;
        const float4 so = wb.shO[si], sd = wb.shD[si];
        r = makeWtRay(xyz(so), xyz(sd));
        tmax = so.w;
        node = 0;
        sp = 0;
        if (r.valid && S.triCount > 0) return true;
        unoccluded(si);       // nothing can block it
        return false;
    };
    for (;;) {
        const int nIdle = __popcll(__ballot(!active));
        if (nIdle >= kRefill || nIdle == 64) {
            if (refill<BATCH>(pool, S, scenes, wb, bounce, kSegShadow + kind, wk, grab, count, active, start) == ~0ull)
                break;
        }
        if (!active) continue;
        const int res = anyStep<COUNT, L, W>(S, r, tmax, node, sp, stk, nodes, tris);
        if (res == 2) unoccluded(si);
        if (res != 0) active = false;
    }
    if (COUNT) {
        wave_add_u64(wb.travCtr + 2, nodes);
        wave_add_u64(wb.travCtr + 3, tris);
    }
}

template <bool COUNT, int L, int W, bool BATCH = false>
__global__ __launch_bounds__(kBlock) void k_connect(DevScene sc, WaveBuffers wb, int bounce, int kind,
                                                     const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    const int nxt = (bounce & 1) ^ 1;
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegShadow + kind) : ctr[kCtrShadow + kind];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    const size_t off = (size_t)kind * (size_t)wb.cap;
    uint32_t nodes = 0, tris = 0;
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkConnect + kind);
    for (int chunk = first_chunk(grab); chunk < count; chunk = next_chunk(wk, grab, count)) {
        if constexpr (!BATCH) {
            for (int row = 0; row < wb.grabRows; ++row) {
                const int si = chunk + row * 64 + lane_id();
                if (si < count) connectOne<COUNT, L, W>(sc, wb, nxt, off + si, stk, &nodes, &tris);
            }
        } else {
            const SegChunk k = segChunk(wb, bounce, kSegShadow + kind, chunk);
            if (k.left <= 0) continue;
            const DevScene S = fetchScene(scenes, k.s);
            for (int row = 0; row < wb.grabRows; ++row) {
                const int e = row * 64 + lane_id();
                if (e < k.left) connectOne<COUNT, L, W>(S, wb, nxt, off + k.first + e, stk, &nodes, &tris);
            }
        }
    }
    if (COUNT) {
        wave_add_u64(wb.travCtr + 2, nodes);
        wave_add_u64(wb.travCtr + 3, tris);
    }
}

// Spotlight pass (rt_direct_lighting.cuh:175-230, rt_optix_shading.cuh:163):
// each request is one path's diffuse-type hit; the lane evaluates every
// spotlight of the scene in index order and traces the shadow ray of each one
// that lights the point, adding unoccluded contributions in that order (the
// megakernel's order), so any number of spotlights costs one record per path.
template <bool COUNT, int L, int WS>
BTE_DEV int spotOne(const DevScene& sc, const WaveBuffers& wb, int nxt, size_t si, size_t qi, const StackRef<L>& stk,
                    uint32_t* nodes, uint32_t* tris) {
    const float4 a = wb.shO[si], b = wb.shD[si], c = wb.shC[si], x = wb.shX[qi];
    Its its;
    its.p = xyz(a);
    its.n = xyz(b);
    its.tri = __float_as_int(b.w);
    Bsdf bsdf = loadBsdf(sc, its.tri);
    applyTextures(sc, its.tri, c.w, x.w, bsdf);
    const Frame f = makeFrame(its.n);
    const uint32_t dest = __float_as_uint(a.w);
    float4* rp = (dest & kFinalBit) ? (wb.finalRad + (dest & ~kFinalBit)) : (wb.aux[nxt] + 2 * (size_t)dest);
    int rays = 0;
    for (int s = 0; s < sc.spotCount; ++s) {
        const NeeRec rec = spotNEE<0>(sc, xyz(c), xyz(x), its, f, bsdf, s);
        if (!rec.ok) continue;
        ++rays;
        float t, u, v;
        if (!traverseW<WS, true, COUNT, L>(sc, its.p, rec.d, RT_EPS, rec.tMax, stk, t, u, v, nodes, tris)) {
            float4 v4 = *rp;
            v4.x = v4.x + rec.c.x;
            v4.y = v4.y + rec.c.y;
            v4.z = v4.z + rec.c.z;
            *rp = v4;
        }
    }
    return rays;
}

template <bool COUNT, int L, int W, bool BATCH = false>
__global__ __launch_bounds__(kBlock) void k_connect_spots(DevScene sc, WaveBuffers wb, int bounce, int kind,
                                                           const DevScene* __restrict__ scenes) {
    constexpr int WS = W >= 4 ? 4 : 2;   // BVH8 scenes carry the BVH4 too
    __shared__ int lds[L * kBlock];
    const int nxt = (bounce & 1) ^ 1;
    int* ctr = wb.counters + (size_t)bounce * kCtrStride;
    const int count = BATCH ? segSlots(wb, bounce, kSegShadow + kind) : ctr[kCtrShadow + kind];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    const size_t off = (size_t)kind * (size_t)wb.cap;
    uint32_t nodes = 0, tris = 0;
    int rays = 0;
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkConnect + kind);
    for (int chunk = first_chunk(grab); chunk < count; chunk = next_chunk(wk, grab, count)) {
        if constexpr (!BATCH) {
            for (int row = 0; row < wb.grabRows; ++row) {
                const int q = chunk + row * 64 + lane_id();
                if (q < count) rays += spotOne<COUNT, L, WS>(sc, wb, nxt, off + q, (size_t)q, stk, &nodes, &tris);
            }
        } else {
            const SegChunk k = segChunk(wb, bounce, kSegShadow + kind, chunk);
            if (k.left <= 0) continue;
            const DevScene S = fetchScene(scenes, k.s);
            for (int row = 0; row < wb.grabRows; ++row) {
                const int e = row * 64 + lane_id();
                if (e < k.left)
                    rays += spotOne<COUNT, L, WS>(S, wb, nxt, off + k.first + e, (size_t)(k.first + e), stk, &nodes,
                                                  &tris);
            }
        }
    }
    unsigned long long r = (unsigned)rays;
    for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane_id() == 0 && r) atomicAdd(ctr + kCtrSpotRays, (int)r);
    if (COUNT) {
        wave_add_u64(wb.travCtr + 2, nodes);
        wave_add_u64(wb.travCtr + 3, tris);
    }
}

// ---------------------------------------------------------------------------
// Tail pass.  Past the first few bounces few paths are left, and a bounce's
// extend / categorise / shade / connect launches each cost more in launch and
// drain time than in work.  From bounce `bounce` on, one persistent kernel
// takes the queued paths and runs each to its end: closest hit, emission,
// next-event estimation with the shadow rays traced at once (area light,
// spotlights in index order, sky: the order in which the connect passes add
// their contributions), scatter and Russian roulette, bounce after bounce.
// A lane whose path has ended takes the next queued path, so lanes of a wave
// may be at different bounces.  Every path is computed by the same functions
// in the same order as in the per-bounce passes, so its radiance, hit
// sequence and AOVs are bit-identical to theirs.
// Ray, hit and shadow-ray counts are kept per bounce in LDS and added to each
// bounce's counter block when the block ends.
// ---------------------------------------------------------------------------
constexpr int kTailStatBounces = 64;   // bounces past the first counted in LDS (later ones: global atomics)
enum { kTsRays, kTsHits, kTsArea, kTsEnv, kTsSpot, kTailStats };

BTE_DEV int tailCounter(const WaveBuffers& wb, int what) {
    switch (what) {
        case kTsRays: return kCtrRays;
        case kTsHits: return kCtrHits;
        case kTsArea: return kCtrShadow + wb.kindArea;
        case kTsEnv: return kCtrShadow + wb.kindEnv;
        default: return kCtrSpotRays;
    }
}

// Closest hit of one ray by a whole traversal (BVH2 or BVH4).
template <int WS, int L>
BTE_DEV int closestHit(const DevScene& S, V3 o, V3 d, const StackRef<L>& stk, float& t, float& u, float& v) {
    return traverseW<WS, false, false, L>(S, o, d, RT_EPS, 1e30f, stk, t, u, v, nullptr, nullptr);
}
template <int WS, int L>
BTE_DEV bool occluded(const DevScene& S, V3 o, V3 d, float tMax, const StackRef<L>& stk) {
    float t, u, v;
    return traverseW<WS, true, false, L>(S, o, d, RT_EPS, tMax, stk, t, u, v, nullptr, nullptr) != 0;
}

#ifndef BTE_TAIL_WAVES
#define BTE_TAIL_WAVES 3   // minimum resident waves per SIMD the register allocation must allow
#endif
template <int L, int WS, bool BATCH = false>
__global__ __launch_bounds__(kBlock, BTE_TAIL_WAVES) void k_tail(DevScene sc, WaveBuffers wb, int bounce, int maxDepth,
                                                  const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    __shared__ int sStat[kTailStatBounces * kTailStats];
    for (int i = threadIdx.x; i < kTailStatBounces * kTailStats; i += kBlock) sStat[i] = 0;
    __syncthreads();
    const int cur = bounce & 1;
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : wb.counters[(size_t)bounce * kCtrStride + kCtrRays];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkTail);
    RayPool pool = makePool(grab);
    DevScene S = sc;
    auto stat = [&](int b, int what) {
        const int j = b - bounce;
        if (j < kTailStatBounces) atomicAdd(&sStat[j * kTailStats + what], 1);
        else atomicAdd(wb.counters + (size_t)b * kCtrStride + tailCounter(wb, what), 1);
    };
    // the lane's path
    bool active = false;
    V3 o = {0, 0, 0}, d = {0, 0, 0}, thr = {0, 0, 0}, rad = {0, 0, 0};
    float brdfPdf = 0.0f, eta = 1.0f;
    uint32_t rng = 0, pid = 0, aovBit = 0;
    bool spec = false;
    int b = bounce;
    auto start = [&](int slot) {
        const float4 ro = wb.rayO[cur][slot], rd = wb.rayD[cur][slot];
        float4 th, ra;
        uint2 st;
        loadAux(wb, cur, slot, th, ra, st);
        o = xyz(ro); brdfPdf = ro.w;
        d = xyz(rd); eta = rd.w;
        thr = xyz(th);
        rad = xyz(ra);
        rng = st.x;
        pid = st.y & kPidMask;
        aovBit = st.y & kAovBit;
        spec = (st.y & kSpecBit) != 0;
        b = bounce;
        return true;
    };
    for (;;) {
        if (__ballot(!active) != 0ull) {
            if (refill<BATCH>(pool, S, scenes, wb, bounce, kSegRays, wk, grab, count, active, start) == ~0ull) break;
        }
        if (!active) continue;
        float t = 0.0f, u = 0.0f, v = 0.0f;
        const int tri = closestHit<WS, L>(S, o, d, stk, t, u, v);
        if (wb.hitCapture && b < wb.maxRec) wb.hitCapture[(size_t)pid * wb.maxRec + b] = tri;
        if (tri < 0) {                                   // as shadeMiss
            envHit(S, d, spec, brdfPdf, thr, rad);
            wb.finalRad[pid] = make_float4(rad.x, rad.y, rad.z, 0.0f);
            active = false;
            continue;
        }
        stat(b, kTsHits);
        const Its its = makeIts(S, o, d, make_float4(__int_as_float(tri), t, u, v));
        Bsdf bsdf = loadBsdf(S, tri);
        applyTextures(S, tri, u, v, bsdf);
        if (S.triFlags && S.triFlags[tri] != 0) emitterHit(S, o, d, spec, brdfPdf, thr, its, rad);
        const Frame frame = makeFrame(its.n);
        if (bsdfIsDiffuse(bsdf.type)) {                  // as shadeRow, then the connect passes in class order
            spec = false;
            const float a0 = rngNext(rng), a1 = rngNext(rng);
            float e0 = 0.0f, e1 = 0.0f;
            if (S.hasEnv) { e0 = rngNext(rng); e1 = rngNext(rng); }
            if (wb.aovAlb && !aovBit) {
                wb.aovAlb[pid] = make_float4(bsdf.p0.x, bsdf.p0.y, bsdf.p0.z, 0.0f);
                wb.aovNrm[pid] = make_float4(its.n.x, its.n.y, its.n.z, 0.0f);
                aovBit = kAovBit;
            }
            NeeRec r = areaNEE<0>(S, d, thr, its, frame, bsdf, a0, a1);
            if (r.ok) {
                stat(b, kTsArea);
                if (!occluded<WS, L>(S, its.p, r.d, r.tMax, stk)) rad = add3(rad, r.c);
            }
            for (int s = 0; s < S.spotCount; ++s) {
                r = spotNEE<0>(S, d, thr, its, frame, bsdf, s);
                if (!r.ok) continue;
                stat(b, kTsSpot);
                if (!occluded<WS, L>(S, its.p, r.d, r.tMax, stk)) rad = add3(rad, r.c);
            }
            if (S.hasEnv) {
                r = envNEE<0>(S, d, thr, its, frame, bsdf, e0, e1);
                if (r.ok) {
                    stat(b, kTsEnv);
                    if (!occluded<WS, L>(S, its.p, r.d, r.tMax, stk)) rad = add3(rad, r.c);
                }
            }
        } else {
            spec = true;
        }
        scatter<0>(its, frame, bsdf, rng, o, d, thr, eta, brdfPdf);
        if (roulette(b, rng, thr, eta) && b + 1 < maxDepth) {
            ++b;
            stat(b, kTsRays);
        } else {
            wb.finalRad[pid] = make_float4(rad.x, rad.y, rad.z, 0.0f);
            active = false;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kTailStatBounces * kTailStats; i += kBlock) {
        const int n = sStat[i];
        if (n) atomicAdd(wb.counters + (size_t)(bounce + i / kTailStats) * kCtrStride + tailCounter(wb, i % kTailStats), n);
    }
}

// Tail pass over a BVH4 or BVH8 as a step machine.  Each lane holds its path
// and one query in flight (the path's closest hit, or one of its shadow rays).
// An iteration advances every lane's query by one node or one triangle: closest
// and shadow queries share one step (children sorted by entry distance in
// both; a shadow ray stops at its first blocker), so lanes of a wave stay
// converged whatever query they trace.  Lanes whose query has finished wait
// until kTailService or more of the wave do (or none is running); then each
// consumes its result (shading after a closest hit, the contribution of an
// unblocked shadow ray) and starts its next query: next-event estimation of
// the hit one light at a time (area light, spotlights in index order, sky),
// then the next bounce.  Finished paths are replaced from the queue.
// Shadow contributions are added in that order, after the emission of the
// hit, as the connect passes add them.
#ifndef BTE_TAIL_SERVICE
#define BTE_TAIL_SERVICE 48   // measured against 8, 16, 32, 40 and 64
#endif
constexpr int kTailService = BTE_TAIL_SERVICE;

// One step of a closest (any = false) or shadow (any = true) query: returns 1
// when a shadow ray is blocked, 2 when the stack is empty (done), else 0.
// tmax is the query's limit (a closest query's best hit so far).
template <int L, int W>
BTE_DEV int queryStep(const DevScene& S, const WtRay& r, bool any, float& tmax, int& node, int& sp,
                      const StackRef<L>& stk, int& bestTri, float& tO, float& uO, float& vO) {
    constexpr float kNone = __builtin_huge_valf();
    if (W == 8 && node >= 0) {
        float k[8];
        int c[8];
        enter8(S, r, node, tmax, k, c);
        sortPair(k[0], c[0], k[2], c[2]); sortPair(k[1], c[1], k[3], c[3]);
        sortPair(k[4], c[4], k[6], c[6]); sortPair(k[5], c[5], k[7], c[7]);
        sortPair(k[0], c[0], k[4], c[4]); sortPair(k[1], c[1], k[5], c[5]);
        sortPair(k[2], c[2], k[6], c[6]); sortPair(k[3], c[3], k[7], c[7]);
        sortPair(k[0], c[0], k[1], c[1]); sortPair(k[2], c[2], k[3], c[3]);
        sortPair(k[4], c[4], k[5], c[5]); sortPair(k[6], c[6], k[7], c[7]);
        sortPair(k[2], c[2], k[4], c[4]); sortPair(k[3], c[3], k[5], c[5]);
        sortPair(k[1], c[1], k[4], c[4]); sortPair(k[3], c[3], k[6], c[6]);
        sortPair(k[1], c[1], k[2], c[2]); sortPair(k[3], c[3], k[4], c[4]); sortPair(k[5], c[5], k[6], c[6]);
        if (k[0] != kNone) {
#pragma unroll
            for (int j = 7; j >= 1; --j)
                if (k[j] != kNone) push(stk, sp, c[j]);
            node = c[0];
            return 0;
        }
    } else if (node >= 0) {
        const float4* np = S.nodes4 + 8 * (size_t)node;
        const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        const int4 ch = *reinterpret_cast<const int4*>(np + 6);
        float f0, f1, f2, f3;
        const float e0 = boxEnter(r, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, RT_EPS, tmax, f0);
        const float e1 = boxEnter(r, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, RT_EPS, tmax, f1);
        const float e2 = boxEnter(r, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, RT_EPS, tmax, f2);
        const float e3 = boxEnter(r, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, RT_EPS, tmax, f3);
        float k0 = e0 <= f0 ? e0 : kNone, k1 = e1 <= f1 ? e1 : kNone;
        float k2 = e2 <= f2 ? e2 : kNone, k3 = e3 <= f3 ? e3 : kNone;
        int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
        sortPair(k0, c0, k1, c1);
        sortPair(k2, c2, k3, c3);
        sortPair(k0, c0, k2, c2);
        sortPair(k1, c1, k3, c3);
        sortPair(k1, c1, k2, c2);
        if (k0 != kNone) {
            push3(stk, sp, c1, k1 != kNone, c2, k2 != kNone, c3, k3 != kNone);
            node = c0;
            return 0;
        }
    } else {
        const int code = ~node;
        const int start = code >> kLeafCountBits;
        const int cnt = (code & (kMaxLeafTris - 1)) + 1;
        const float4* tp = S.leafTris + 3 * (size_t)start;
        const float4 a = tp[0], b = tp[1], c = tp[2];
        float t, u, v;
        if (hitTriangle(r, a, b, c, RT_EPS, tmax, t, u, v)) {
            if (any) return 1;
            const int id = __float_as_int(a.w);
            if (t < tmax || bestTri < 0 || id < bestTri) {
                tmax = t; bestTri = id; tO = t; uO = u; vO = v;
            }
        }
        if (cnt > 1) {   // the rest of the leaf is the next step
            node = ~(((start + 1) << kLeafCountBits) | (cnt - 2));
            return 0;
        }
    }
    if (sp == 0) return 2;
    node = pop(stk, sp);
    return 0;
}

template <int L, int W, bool BATCH = false>
__global__ __launch_bounds__(kBlock, BTE_TAIL_WAVES) void k_tail_steps(DevScene sc, WaveBuffers wb, int bounce,
                                                                        int maxDepth,
                                                                        const DevScene* __restrict__ scenes) {
    __shared__ int lds[L * kBlock];
    __shared__ int sStat[kTailStatBounces * kTailStats];
    for (int i = threadIdx.x; i < kTailStatBounces * kTailStats; i += kBlock) sStat[i] = 0;
    __syncthreads();
    const int cur = bounce & 1;
    const int count = BATCH ? segSlots(wb, bounce, kSegRays) : wb.counters[(size_t)bounce * kCtrStride + kCtrRays];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<L> stk{lds + threadIdx.x, wb.ovfStack + (size_t)gtid * kOvfStack};
    const int grab = 64 * wb.grabRows;
    int* wk = work_group(wb, bounce, kWorkTail);
    RayPool pool = makePool(grab);
    DevScene S = sc;
    auto stat = [&](int b, int what) __attribute__((always_inline)) {
        const int j = b - bounce;
        if (j < kTailStatBounces) atomicAdd(&sStat[j * kTailStats + what], 1);
        else atomicAdd(wb.counters + (size_t)b * kCtrStride + tailCounter(wb, what), 1);
    };
    // the lane's path: o, d = its next extension ray
    bool hasPath = false, spec = false, survive = false;
    V3 o = {0, 0, 0}, d = {0, 0, 0}, thr = {0, 0, 0}, rad = {0, 0, 0};
    float brdfPdf = 0.0f, eta = 1.0f;
    uint32_t rng = 0, pid = 0, aovBit = 0;
    int b = bounce;
    // the last hit, for its next-event estimation: stages [stage, stageEnd) =
    // area light, spotlights 0..spotCount-1, sky
    V3 hp = {0, 0, 0}, hn = {0, 0, 0}, dIn = {0, 0, 0}, thrIn = {0, 0, 0};
    int hTri = 0, stage = 0, stageEnd = 0;
    float hu = 0.0f, hv = 0.0f, a0 = 0.0f, a1 = 0.0f, e0 = 0.0f, e1 = 0.0f;
    // the query in flight
    bool running = false, shadow = false;
    WtRay r = makeWtRay(V3{0, 0, 0}, V3{0, 0, 1});
    float qmax = 0.0f, qt = 0.0f, qu = 0.0f, qv = 0.0f;
    int node = 0, sp = 0, qTri = -1, qEnd = 0;
    V3 pendC = {0, 0, 0};

    // Values a later step cannot read are reset where they die, so that the
    // compiler does not keep them in registers across the loop.
    auto endQuery = [&]() __attribute__((always_inline)) {
        r = makeWtRay(V3{0, 0, 0}, V3{0, 0, 1});
        qmax = qt = qu = qv = 0.0f;
        node = sp = 0;
        qTri = -1;
        pendC = V3{0, 0, 0};
    };
    auto endHit = [&]() __attribute__((always_inline)) {
        hp = hn = dIn = thrIn = V3{0, 0, 0};
        hTri = stage = stageEnd = 0;
        hu = hv = a0 = a1 = e0 = e1 = 0.0f;
    };
    auto finish = [&]() __attribute__((always_inline)) {
        wb.finalRad[pid] = make_float4(rad.x, rad.y, rad.z, 0.0f);
        hasPath = false;
        endQuery();
        endHit();
        o = d = thr = rad = V3{0, 0, 0};
    };
    auto miss = [&]() __attribute__((always_inline)) {                               // as shadeMiss
        if (wb.hitCapture && b < wb.maxRec) wb.hitCapture[(size_t)pid * wb.maxRec + b] = -1;
        envHit(S, r.d, spec, brdfPdf, thr, rad);
        finish();
    };
    // Starts the closest query of ray (o, d); an untraceable ray is a miss at once.
    auto startClosest = [&]() __attribute__((always_inline)) {
        r = makeWtRay(o, d);
        o = d = V3{0, 0, 0};                          // the query holds the ray now
        if (!r.valid || S.triCount == 0) {
            miss();
            return;
        }
        qmax = 1e30f;
        node = sp = 0;
        qTri = -1;
        qt = qu = qv = 0.0f;
        shadow = false;
        running = true;
    };
    // Next query of the path: the hit's remaining light stages, then the next bounce.
    auto advance = [&]() __attribute__((always_inline)) {
        while (stage < stageEnd) {
            const int k = stage++;
            Its its;
            its.p = hp; its.n = hn; its.tri = hTri;
            Bsdf bsdf = loadBsdf(S, hTri);
            applyTextures(S, hTri, hu, hv, bsdf);
            const Frame f = makeFrame(hn);
            NeeRec rec;
            rec.ok = false;
            int what = kTsArea;
            if (k == 0) {
                rec = areaNEE<0>(S, dIn, thrIn, its, f, bsdf, a0, a1);
            } else if (k <= S.spotCount) {
                rec = spotNEE<0>(S, dIn, thrIn, its, f, bsdf, k - 1);
                what = kTsSpot;
            } else if (S.hasEnv) {
                rec = envNEE<0>(S, dIn, thrIn, its, f, bsdf, e0, e1);
                what = kTsEnv;
            }
            if (!rec.ok) continue;
            stat(b, what);
            r = makeWtRay(hp, rec.d);
            if (!r.valid || S.triCount == 0) {       // nothing can block it
                rad = add3(rad, rec.c);
                continue;
            }
            qmax = rec.tMax;
            node = sp = 0;
            pendC = rec.c;
            shadow = true;
            running = true;
            return;
        }
        if (!survive) {
            finish();
            return;
        }
        endHit();
        ++b;
        stat(b, kTsRays);
        startClosest();
    };
    // Shading after a closest hit (a miss ends the path).
    auto shade = [&]() __attribute__((always_inline)) {
        const int tri = qTri;
        if (tri < 0) {
            miss();
            return;
        }
        if (wb.hitCapture && b < wb.maxRec) wb.hitCapture[(size_t)pid * wb.maxRec + b] = tri;
        stat(b, kTsHits);
        const V3 ro = r.o, rd = r.d;
        const Its its = makeIts(S, ro, rd, make_float4(__int_as_float(tri), qt, qu, qv));
        Bsdf bsdf = loadBsdf(S, tri);
        applyTextures(S, tri, qu, qv, bsdf);
        if (S.triFlags && S.triFlags[tri] != 0) emitterHit(S, ro, rd, spec, brdfPdf, thr, its, rad);
        const Frame frame = makeFrame(its.n);
        stage = stageEnd = 0;
        if (bsdfIsDiffuse(bsdf.type)) {
            spec = false;
            a0 = rngNext(rng);
            a1 = rngNext(rng);
            if (S.hasEnv) { e0 = rngNext(rng); e1 = rngNext(rng); }
            if (wb.aovAlb && !aovBit) {
                wb.aovAlb[pid] = make_float4(bsdf.p0.x, bsdf.p0.y, bsdf.p0.z, 0.0f);
                wb.aovNrm[pid] = make_float4(its.n.x, its.n.y, its.n.z, 0.0f);
                aovBit = kAovBit;
            }
            hp = its.p; hn = its.n; hTri = tri; hu = qu; hv = qv;
            dIn = rd; thrIn = thr;
            stageEnd = S.spotCount + 2;
        } else {
            spec = true;
        }
        d = rd;
        scatter<0>(its, frame, bsdf, rng, o, d, thr, eta, brdfPdf);
        survive = roulette(b, rng, thr, eta) && b + 1 < maxDepth;
    };
    auto start = [&](int slot) __attribute__((always_inline)) {
        const float4 ro = wb.rayO[cur][slot], rd = wb.rayD[cur][slot];
        float4 th, ra;
        uint2 st;
        loadAux(wb, cur, slot, th, ra, st);
        o = xyz(ro); brdfPdf = ro.w;
        d = xyz(rd); eta = rd.w;
        thr = xyz(th);
        rad = xyz(ra);
        rng = st.x;
        pid = st.y & kPidMask;
        aovBit = st.y & kAovBit;
        spec = (st.y & kSpecBit) != 0;
        b = bounce;
        hasPath = true;
        running = false;
        startClosest();
        return hasPath;
    };

    for (;;) {
        const unsigned long long busy = __ballot(running);
        if (64 - __popcll(busy) >= kTailService || busy == 0ull) {
            if (hasPath && !running) {
                if (!shadow) shade();
                else if (qEnd == 2) rad = add3(rad, pendC);   // an unblocked shadow ray
                if (hasPath) advance();
            }
            if (__ballot(!hasPath) != 0ull) {
                if (refill<BATCH>(pool, S, scenes, wb, bounce, kSegRays, wk, grab, count, hasPath, start) == ~0ull)
                    break;
            }
        }
        if (running) {
            qEnd = queryStep<L, W>(S, r, shadow, qmax, node, sp, stk, qTri, qt, qu, qv);
            if (qEnd != 0) running = false;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kTailStatBounces * kTailStats; i += kBlock) {
        const int n = sStat[i];
        if (n) atomicAdd(wb.counters + (size_t)(bounce + i / kTailStats) * kCtrStride + tailCounter(wb, i % kTailStats), n);
    }
}

// Welford update of one pixel's running means with `frames` new samples: path
// ids src + f * stride, output pixel `out`.
BTE_DEV void presentOne(const WaveBuffers& wb, int frames, int frameBase, size_t src, size_t stride, size_t out,
                        float4* accum, double* sum64, uint32_t* fb) {
    if (wb.aovAlb) {
        // running means of the AOVs, same Welford update as the radiance
        float4 ma = wb.aovAccAlb[out], mn = wb.aovAccNrm[out];
        for (int f = 0; f < frames; ++f) {
            const float4 a = wb.aovAlb[src + (size_t)f * stride], nr = wb.aovNrm[src + (size_t)f * stride];
            const float n = (float)((uint32_t)(frameBase + f) + 1u);
            ma.x = ma.x + (a.x - ma.x) / n; ma.y = ma.y + (a.y - ma.y) / n; ma.z = ma.z + (a.z - ma.z) / n;
            mn.x = mn.x + (nr.x - mn.x) / n; mn.y = mn.y + (nr.y - mn.y) / n; mn.z = mn.z + (nr.z - mn.z) / n;
        }
        wb.aovAccAlb[out] = ma;
        wb.aovAccNrm[out] = mn;
    }
    float4 mean = accum[out];
    // binary64 sums take one frame at a time, in frame order, so they do not
    // depend on how the frames were split into passes
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (sum64) {
        sx = sum64[3 * out];
        sy = sum64[3 * out + 1];
        sz = sum64[3 * out + 2];
    }
    for (int f = 0; f < frames; ++f) {
        const float4 x = wb.finalRad[src + (size_t)f * stride];
        const float n = (float)((uint32_t)(frameBase + f) + 1u);
        mean.x = mean.x + (x.x - mean.x) / n;
        mean.y = mean.y + (x.y - mean.y) / n;
        mean.z = mean.z + (x.z - mean.z) / n;
        sx += (double)x.x; sy += (double)x.y; sz += (double)x.z;
    }
    accum[out] = mean;
    if (sum64) {
        sum64[3 * out] = sx;
        sum64[3 * out + 1] = sy;
        sum64[3 * out + 2] = sz;
    }
    if (fb) fb[out] = packGamma(V3{mean.x, mean.y, mean.z});
}

__global__ __launch_bounds__(kBlock) void k_present(WaveBuffers wb, int pixels, int frames, float4* accum,
                                                     double* sum64, uint32_t* fb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    presentOne(wb, frames, *wb.frameBase, (size_t)p, (size_t)pixels, (size_t)p, accum, sum64, fb);
}

// Batch: blockIdx.y = scene; its pixels land at its pixelBase.
__global__ __launch_bounds__(kBlock) void k_present_batch(WaveBuffers wb, int frames, float4* accum, double* sum64,
                                                           uint32_t* fb) {
    const BatchSlot b = wb.slots[blockIdx.y];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.pixels) return;
    presentOne(wb, frames, *wb.frameBase, (size_t)(b.pathBase + p), (size_t)b.pixels, (size_t)(b.pixelBase + p),
               accum, sum64, fb);
}

// Megakernel (src/optix_programs.cu:54-108): one pixel per lane, spp paths
// sharing one RNG stream, inline visibility.
__global__ __launch_bounds__(kBlock) void k_megakernel(DevScene sc, int spp, int maxDepth, int frameIndex,
                                                        float4* accum, double* sum64, uint32_t* fb, int* ovf,
                                                        int* work) {
    __shared__ int lds[kLdsStack * kBlock];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<kLdsStack> stk{lds + threadIdx.x, ovf + (size_t)gtid * kOvfStack};
    const int pixels = sc.width * sc.height;
    for (;;) {
        const int base = wave_grab(work);
        if (base >= pixels) break;
        const int p = base + lane_id();
        if (p >= pixels) continue;
        const int px = p % sc.width, py = p / sc.width;
        uint32_t rng = seedRng(hashUint((uint32_t)p ^ ((uint32_t)frameIndex * 0x9e3779b9u)));
        V3 fin = {0, 0, 0};
        for (int i = 0; i < spp; ++i) {
            const float jx = rngNext(rng) - 0.5f;
            const float jy = rngNext(rng) - 0.5f;
            V3 o = {sc.cam.origin.x, sc.cam.origin.y, sc.cam.origin.z};
            V3 d = primaryDir(sc.cam, px, py, sc.width, sc.height, jx, jy);
            V3 thr = {1, 1, 1}, rad = {0, 0, 0};
            float eta = 1.0f, brdfPdf = 0.0f;
            bool spec = true;
            for (int bounce = 0; bounce < maxDepth; ++bounce) {
                float t = 0.0f, u = 0.0f, v = 0.0f;
                const int tri = traverse<false>(sc, o, d, RT_EPS, 1e30f, stk, t, u, v);
                if (tri < 0) { envHit(sc, d, spec, brdfPdf, thr, rad); break; }
                const Its its = makeIts(sc, o, d, make_float4(__int_as_float(tri), t, u, v));
                Bsdf bsdf = loadBsdf(sc, tri);
                applyTextures(sc, tri, u, v, bsdf);
                if (sc.triFlags && sc.triFlags[tri] != 0) emitterHit(sc, o, d, spec, brdfPdf, thr, its, rad);
                const Frame frame = makeFrame(its.n);
                if (bsdfIsDiffuse(bsdf.type)) {
                    spec = false;
                    const float a0 = rngNext(rng), a1 = rngNext(rng);
                    NeeRec r = areaNEE<0>(sc, d, thr, its, frame, bsdf, a0, a1);
                    if (r.ok && !traverse<true>(sc, its.p, r.d, RT_EPS, r.tMax, stk, t, u, v)) rad = add3(rad, r.c);
                    for (int s = 0; s < sc.spotCount; ++s) {
                        r = spotNEE<0>(sc, d, thr, its, frame, bsdf, s);
                        if (r.ok && !traverse<true>(sc, its.p, r.d, RT_EPS, r.tMax, stk, t, u, v)) rad = add3(rad, r.c);
                    }
                    if (sc.hasEnv) {
                        const float e0 = rngNext(rng), e1 = rngNext(rng);
                        r = envNEE<0>(sc, d, thr, its, frame, bsdf, e0, e1);
                        if (r.ok && !traverse<true>(sc, its.p, r.d, RT_EPS, r.tMax, stk, t, u, v)) rad = add3(rad, r.c);
                    }
                } else {
                    spec = true;
                }
                scatter<0>(its, frame, bsdf, rng, o, d, thr, eta, brdfPdf);
                if (!roulette(bounce, rng, thr, eta)) break;
            }
            fin = add3(fin, rad);
        }
        fin = div3(fin, (float)spp);
        float4 mean = accum[p];
        const float n = (float)((uint32_t)frameIndex + 1u);
        mean.x = mean.x + (fin.x - mean.x) / n;
        mean.y = mean.y + (fin.y - mean.y) / n;
        mean.z = mean.z + (fin.z - mean.z) / n;
        accum[p] = mean;
        if (sum64) {
            sum64[3 * (size_t)p] += (double)fin.x;
            sum64[3 * (size_t)p + 1] += (double)fin.y;
            sum64[3 * (size_t)p + 2] += (double)fin.z;
        }
        if (fb) fb[p] = packGamma(V3{mean.x, mean.y, mean.z});
    }
}

// OVF: overflow-stack entries per thread beyond the kLdsStack in LDS (the
// BVH2 traversal holds at most one entry per tree level).
template <int OVF>
__global__ __launch_bounds__(kBlock) void k_trace_rays(DevScene sc, const float* rays, const float* tmaxPer,
                                                        long long n, float tmin, float tmax, int anyHit, int* tri,
                                                        float* tt, float* uu, float* vv, int* ovf, int* work) {
    __shared__ int lds[kLdsStack * kBlock];
    const int gtid = blockIdx.x * kBlock + threadIdx.x;
    const StackRef<kLdsStack> stk{lds + threadIdx.x, ovf + (size_t)gtid * OVF};
    for (;;) {
        const int base = wave_grab(work);
        if (base >= n) break;
        const long long i = base + lane_id();
        if (i >= n) continue;
        const float* r = rays + 6 * i;
        const V3 o = {r[0], r[1], r[2]}, d = {r[3], r[4], r[5]};
        float t = 0.0f, u = 0.0f, v = 0.0f;
        if (anyHit) {
            tri[i] = traverse<true>(sc, o, d, tmin, tmaxPer ? tmaxPer[i] : tmax, stk, t, u, v);
        } else {
            tri[i] = traverse<false>(sc, o, d, tmin, tmax, stk, t, u, v);
            tt[i] = t; uu[i] = u; vv[i] = v;
        }
    }
}

// Env-map CDF (src/host/scene_upload.cu:428-459) in binary64: one block per
// row scans lum(texel) * sin(theta_row), then one block scans the row sums.
constexpr int kScanBlock = 256;
__device__ double block_exclusive_scan(double v, double* sh, double& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const double a = (t >= off) ? sh[t - off] : 0.0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[kScanBlock - 1];
    const double incl = sh[t];
    __syncthreads();
    return incl - v;
}
__global__ __launch_bounds__(kScanBlock) void k_env_rows(const float4* env, int W, int H, float* cond,
                                                          double* rowSums) {
    __shared__ double sh[kScanBlock];
    const int iy = blockIdx.x;
    const double theta = 3.14159265358979323846 * ((double)iy + 0.5) / (double)H;
    const double st = sin(theta);
    const int per = (W + kScanBlock - 1) / kScanBlock;
    const int x0 = threadIdx.x * per;
    double local = 0.0;
    for (int k = 0; k < per; ++k) {
        const int ix = x0 + k;
        if (ix < W) {
            const float4 p = env[(size_t)iy * W + ix];
            local += (0.2126 * (double)p.x + 0.7152 * (double)p.y + 0.0722 * (double)p.z) * st;
        }
    }
    double total;
    double run = block_exclusive_scan(local, sh, total);
    float* row = cond + (size_t)iy * (size_t)(W + 1);
    if (threadIdx.x == 0) row[0] = 0.0f;
    for (int k = 0; k < per; ++k) {
        const int ix = x0 + k;
        if (ix < W) {
            const float4 p = env[(size_t)iy * W + ix];
            run += (0.2126 * (double)p.x + 0.7152 * (double)p.y + 0.0722 * (double)p.z) * st;
            row[ix + 1] = (float)run;
        }
    }
    if (threadIdx.x == 0) rowSums[iy] = total;
}
__global__ __launch_bounds__(kScanBlock) void k_env_marginal(const double* rowSums, int H, float* marg,
                                                              float* total) {
    __shared__ double sh[kScanBlock];
    const int per = (H + kScanBlock - 1) / kScanBlock;
    const int y0 = threadIdx.x * per;
    double local = 0.0;
    for (int k = 0; k < per; ++k)
        if (y0 + k < H) local += rowSums[y0 + k];
    double tot;
    double run = block_exclusive_scan(local, sh, tot);
    if (threadIdx.x == 0) marg[0] = 0.0f;
    for (int k = 0; k < per; ++k) {
        if (y0 + k < H) {
            run += rowSums[y0 + k];
            marg[y0 + k + 1] = (float)run;
        }
    }
    if (threadIdx.x == 0) *total = (float)tot;
}

// Guide tables for cdfSearchGuided: guide[g] = last index i < count with
// cdf[i] <= sum * g / G, g = 0..G (G = count).  Clears *monotone when a CDF
// decreases anywhere or holds a NaN (guides are then not used).
__global__ __launch_bounds__(256) void k_env_guides(const float* marg, const float* cond, int W, int H,
                                                     int* margGuide, int* condGuide, int* monotone) {
    const int r = blockIdx.x;
    const bool isMarg = (r == H);
    const int count = isMarg ? H : W;
    const float* cdf = isMarg ? marg : cond + (size_t)r * (size_t)(W + 1);
    int* guide = isMarg ? margGuide : condGuide + (size_t)r * (size_t)(W + 1);
    const float sum = cdf[count];
    for (int g = threadIdx.x; g <= count; g += blockDim.x) {
        const float thr = sum * ((float)g / (float)count);
        guide[g] = cdfSearch(cdf, count, thr);
    }
    for (int i = threadIdx.x; i < count; i += blockDim.x)
        if (!(cdf[i + 1] >= cdf[i])) atomicAnd(monotone, 0);   // NaN clears it too
}

// Batch: row starts of queues [q0, q0 + n) of one bounce from their per-scene
// counts (one thread per queue); the totals of the ray and shadow queues also
// go to the plain counters the host reads.
__global__ void k_segment_rows(WaveBuffers wb, int bounce, int q0, int n) {
    if ((int)threadIdx.x >= n) return;
    const int q = q0 + (int)threadIdx.x;
    const int* cnt = segCounts(wb, bounce, q);
    int* start = segRowStart(wb, bounce, q);
    int rows = 0, sum = 0;
    for (int s = 0; s < wb.sceneCount; ++s) {
        start[s] = rows;
        const int c = cnt[s];
        sum += c;
        rows += (c + 64 * kSegRowAlign - 1) / (64 * kSegRowAlign) * kSegRowAlign;
    }
    start[wb.sceneCount] = rows;
    if (q == kSegRays) wb.counters[(size_t)bounce * kCtrStride + kCtrRays] = sum;
    if (q >= kSegShadow) wb.counters[(size_t)bounce * kCtrStride + kCtrShadow + (q - kSegShadow)] = sum;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int blocksFor(long long n, int b) { return (int)((n + b - 1) / b); }

void launch_fill_i32(int* p, long long n, int value, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocksFor(n, 256)), dim3(256), 0, s, p, n, value);
}

void launch_generate(const DevScene& sc, const WaveBuffers& wb, long long paths, int pixels, hipStream_t s) {
    if (wb.camPaths > 0 && wb.camRays && !wb.aovAlb) {   // nothing to store: the pass's ray count only
        launch_fill_i32(wb.counters + kCtrRays, 1, (int)paths, s);
        return;
    }
    hipLaunchKernelGGL(k_generate, dim3(blocksFor(paths, kBlock)), dim3(kBlock), 0, s, sc, wb, paths, pixels);
}

// One category's shading kernel, if the scene uses that BSDF type.
template <int T, bool B>
static void launch_shade_cat(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                             hipStream_t s, const DevScene* scenes, bool cam) {
    if (!(wb.catMask & (1 << T))) return;
    if constexpr (!B) {
        if (cam) {
            hipLaunchKernelGGL((k_shade<T, false, true>), dim3(ls.shadeCamBlocks[T]), dim3(kBlock), 0, s, sc, wb, bounce,
                               maxDepth, scenes);
            return;
        }
    }
    hipLaunchKernelGGL((k_shade<T, B>), dim3(ls.shadeCatBlocks[B][T]), dim3(kBlock), 0, s, sc, wb, bounce, maxDepth,
                       scenes);
}

template <int L, int W, bool B>
static void launch_bounce_l(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth,
                            const LaunchShape& ls, hipStream_t s, hipEvent_t* ev, const DevScene* scenes) {
    // traversal counters are a single-scene instrument
    const bool count = !B && wb.travCtr != nullptr;
    if (ev) (void)hipEventRecord(ev[0], s);
    if constexpr (W >= 4) {
        if (count) hipLaunchKernelGGL((k_extend_pool<true, L, false, W>), dim3(ls.extendBlocks), dim3(kBlock), 0, s, sc, wb, bounce, scenes);
        else if (W == 4 && bounce == 0 && ls.packetBlocks > 0)
            hipLaunchKernelGGL(k_extend_packet<B>, dim3(B ? ls.packetBlocksBatch : ls.packetBlocks), dim3(kBlock), 0, s, sc, wb,
                               bounce, scenes);
        else hipLaunchKernelGGL((k_extend_pool<false, L, B, W>), dim3(ls.extendBlocks), dim3(kBlock), 0, s, sc, wb, bounce, scenes);
    } else if (count) {
        hipLaunchKernelGGL((k_extend<true, L, W>), dim3(ls.extendBlocks), dim3(kBlock), 0, s, sc, wb, bounce, scenes);
    } else {
        hipLaunchKernelGGL((k_extend<false, L, W, B>), dim3(ls.extendBlocks), dim3(kBlock), 0, s, sc, wb, bounce,
                           scenes);
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    // bounce 0 of a single scene: the camera-path instantiations (the state is computed, not loaded)
    const bool cam = !B && bounce == 0 && wb.camPaths > 0;
    if constexpr (!B) {
        if (cam) hipLaunchKernelGGL((k_categorize<false, true>), dim3(ls.shadeCamBlocks[kCatMiss]), dim3(kBlock), 0, s, sc, wb, bounce, scenes);
    }
    if (!cam) hipLaunchKernelGGL(k_categorize<B>, dim3(ls.shadeCatBlocks[B][kCatMiss]), dim3(kBlock), 0, s, sc, wb, bounce, scenes);
    if (B) launch_segment_rows(wb, bounce, kSegCat, kNumCats, s);
    // a BSDF category is shaded only when the scene uses that type
    launch_shade_cat<BTE_BSDF_DIFFUSE, B>(sc, wb, bounce, maxDepth, ls, s, scenes, cam);
    launch_shade_cat<BTE_BSDF_DIELECTRIC, B>(sc, wb, bounce, maxDepth, ls, s, scenes, cam);
    launch_shade_cat<BTE_BSDF_MIRROR, B>(sc, wb, bounce, maxDepth, ls, s, scenes, cam);
    launch_shade_cat<BTE_BSDF_MICROFACET, B>(sc, wb, bounce, maxDepth, ls, s, scenes, cam);
    launch_shade_cat<BTE_BSDF_DISNEY, B>(sc, wb, bounce, maxDepth, ls, s, scenes, cam);
    if (B) {
        launch_segment_rows(wb, bounce + 1, kSegRays, 1, s);
        launch_segment_rows(wb, bounce, kSegShadow, wb.kinds, s);
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    for (int k = 0; k < wb.kinds; ++k) {
        if (k == wb.kindSpot) {
            if (count)
                hipLaunchKernelGGL((k_connect_spots<true, L, W>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc, wb,
                                   bounce, k, scenes);
            else
                hipLaunchKernelGGL((k_connect_spots<false, L, W, B>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc,
                                   wb, bounce, k, scenes);
            continue;
        }
        if constexpr (W >= 4) {
            if (count)
                hipLaunchKernelGGL((k_connect_pool<true, L, false, W>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc, wb, bounce, k, scenes);
            else
                hipLaunchKernelGGL((k_connect_pool<false, L, B, W>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc, wb, bounce, k, scenes);
        } else if (count) {
            hipLaunchKernelGGL((k_connect<true, L, W>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc, wb, bounce, k,
                               scenes);
        } else {
            hipLaunchKernelGGL((k_connect<false, L, W, B>), dim3(ls.connectBlocks), dim3(kBlock), 0, s, sc, wb,
                               bounce, k, scenes);
        }
    }
    if (ev) (void)hipEventRecord(ev[3], s);
}

template <bool B, int L>
static void launch_bounce_w(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth,
                            const LaunchShape& ls, hipStream_t s, hipEvent_t* ev, const DevScene* scenes, int width) {
    if (width == 8) launch_bounce_l<L, 8, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
    else if (width == 4) launch_bounce_l<L, 4, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
    else launch_bounce_l<L, 2, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
}

template <bool B>
static void launch_bounce_b(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth,
                            const LaunchShape& ls, hipStream_t s, hipEvent_t* ev, const DevScene* scenes, int width) {
    switch (ls.ldsStack) {
        case 8: launch_bounce_w<B, 8>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
        case 16: launch_bounce_w<B, 16>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
        default: launch_bounce_w<B, 32>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
    }
}

void launch_bounce(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                   hipStream_t s, hipEvent_t* ev, const DevScene* scenes) {
    // the host checks that every scene of a batch carries the chosen layout
    int width = ls.bvhWidth;
    if (!scenes) {
        if (width == 8 && sc.nodes8 == nullptr) width = 4;
        if (width == 4 && sc.nodes4 == nullptr) width = 2;
    }
    if (scenes) launch_bounce_b<true>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width);
    else launch_bounce_b<false>(sc, wb, bounce, maxDepth, ls, s, ev, nullptr, width);
}

template <int L, int W, bool B>
static void launch_tail_l(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                          hipStream_t s, hipEvent_t* ev, const DevScene* scenes) {
    if (ev) (void)hipEventRecord(ev[0], s);
    if constexpr (W >= 4)
        hipLaunchKernelGGL((k_tail_steps<L, W, B>), dim3(ls.tailBlocks), dim3(kBlock), 0, s, sc, wb, bounce, maxDepth,
                           scenes);
    else
        hipLaunchKernelGGL((k_tail<L, 2, B>), dim3(ls.tailBlocks), dim3(kBlock), 0, s, sc, wb, bounce, maxDepth,
                           scenes);
    if (ev)
        for (int i = 1; i < 4; ++i) (void)hipEventRecord(ev[i], s);
}

template <bool B, int L>
static void launch_tail_w(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                          hipStream_t s, hipEvent_t* ev, const DevScene* scenes, int width) {
    if (width == 8) launch_tail_l<L, 8, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
    else if (width == 4) launch_tail_l<L, 4, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
    else launch_tail_l<L, 2, B>(sc, wb, bounce, maxDepth, ls, s, ev, scenes);
}

template <bool B>
static void launch_tail_b(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                          hipStream_t s, hipEvent_t* ev, const DevScene* scenes, int width) {
    switch (ls.ldsStack) {
        case 8: launch_tail_w<B, 8>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
        case 16: launch_tail_w<B, 16>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
        default: launch_tail_w<B, 32>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width); break;
    }
}

void launch_tail(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                 hipStream_t s, hipEvent_t* ev, const DevScene* scenes) {
    // the host checks that every scene of a batch carries the chosen layout
    int width = ls.bvhWidth;
    if (!scenes) {
        if (width == 8 && sc.nodes8 == nullptr) width = 4;
        if (width == 4 && sc.nodes4 == nullptr) width = 2;
    }
    if (scenes) launch_tail_b<true>(sc, wb, bounce, maxDepth, ls, s, ev, scenes, width);
    else launch_tail_b<false>(sc, wb, bounce, maxDepth, ls, s, ev, nullptr, width);
}

void launch_generate_batch(const DevScene* scenes, const WaveBuffers& wb, int maxScenePaths, long long totalSlots,
                           hipStream_t s) {
    if (maxScenePaths <= 0 || wb.sceneCount <= 0) return;
    hipLaunchKernelGGL(k_generate_batch, dim3(blocksFor(maxScenePaths, kBlock), wb.sceneCount), dim3(kBlock), 0, s,
                       scenes, wb, totalSlots);
    launch_segment_rows(wb, 0, kSegRays, 1, s);
}

void launch_segment_rows(const WaveBuffers& wb, int bounce, int q, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_segment_rows, dim3(1), dim3(64), 0, s, wb, bounce, q, n);
}

void launch_present_batch(const WaveBuffers& wb, int maxScenePixels, int frames, float4* accum, double* sum64,
                          uint32_t* fb, hipStream_t s) {
    if (maxScenePixels <= 0 || wb.sceneCount <= 0) return;
    hipLaunchKernelGGL(k_present_batch, dim3(blocksFor(maxScenePixels, kBlock), wb.sceneCount), dim3(kBlock), 0, s,
                       wb, frames, accum, sum64, fb);
}

void launch_present(const DevScene& sc, const WaveBuffers& wb, int pixels, int frames, float4* accum,
                    double* sum64, uint32_t* fb, hipStream_t s) {
    (void)sc;
    hipLaunchKernelGGL(k_present, dim3(blocksFor(pixels, kBlock)), dim3(kBlock), 0, s, wb, pixels, frames, accum,
                       sum64, fb);
}

void launch_megakernel(const DevScene& sc, int spp, int maxDepth, int frameIndex, float4* accum, double* sum64,
                       uint32_t* fb, int* ovfStack, hipStream_t s) {
    // work counter lives in the last int of the overflow area's header
    int* work = ovfStack;
    int* ovf = ovfStack + 64;
    launch_fill_i32(work, 1, 0, s);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const LaunchShape ls = query_launch_shape(dev, 32, 0, 4, 2);
    hipLaunchKernelGGL(k_megakernel, dim3(ls.auxBlocks), dim3(kBlock), 0, s, sc, spp, maxDepth, frameIndex, accum,
                       sum64, fb, ovf, work);
}

void launch_trace_rays(const DevScene& sc, const float* rays, const float* tmaxPerRay, long long n, float tmin,
                       float tmax, int anyHit, int* tri, float* t, float* u, float* v, int* ovfStack, hipStream_t s) {
    int* work = ovfStack;
    int* ovf = ovfStack + 64;
    launch_fill_i32(work, 1, 0, s);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const LaunchShape ls = query_launch_shape(dev, 32, 0, 4, 2);
    hipLaunchKernelGGL(k_trace_rays<kOvfStack>, dim3(ls.auxBlocks), dim3(kBlock), 0, s, sc, rays, tmaxPerRay, n, tmin, tmax,
                       anyHit, tri, t, u, v, ovf, work);
}

// ---------------------------------------------------------------------------
// Allocation-free ray tracing: the work counter and the traversal stacks'
// overflow live in one static device array, `scratch`, of a fixed size that
// every device holds for the life of the process.  Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): the address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr int kNoAllocBlocks = 512;            // persistent grid (256 threads per block)
constexpr int kNoAllocOvf = 32;                // overflow entries per thread beyond the 32 in LDS
constexpr int kNoAllocHeader = 64;             // ints before the overflow area; int 0 = work counter
constexpr size_t kNoAllocScratchInts = (size_t)kNoAllocHeader + (size_t)kNoAllocBlocks * kBlock * kNoAllocOvf;
static_assert(kNoAllocScratchInts * sizeof(int) <= (16u << 20) + 256u, "scratch stays at 16 MiB + 256 B");
__device__ __align__(256) int scratch[kNoAllocScratchInts];

constexpr int kMaxScratchDevices = 64;
static int* g_scratchDev[kMaxScratchDevices] = {};

hipError_t noalloc_scratch(int** ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
    if (!g_scratchDev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchDev[dev] = static_cast<int*>(p);
    }
    *ptr = g_scratchDev[dev];
    return hipSuccess;
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchInts * sizeof(int); }
int noalloc_max_bvh_depth() { return kLdsStack + kNoAllocOvf; }

void launch_trace_rays_noalloc(const DevScene& sc, const float* rays, const float* tmaxPerRay, long long n,
                               float tmin, float tmax, int anyHit, int* tri, float* t, float* u, float* v,
                               int* scratchPtr, hipStream_t s) {
    launch_fill_i32(scratchPtr, 1, 0, s);   // work counter (a kernel, not a memset node)
    hipLaunchKernelGGL(k_trace_rays<kNoAllocOvf>, dim3(kNoAllocBlocks), dim3(kBlock), 0, s, sc, rays, tmaxPerRay, n,
                       tmin, tmax, anyHit, tri, t, u, v, scratchPtr + kNoAllocHeader, scratchPtr);
}

void launch_env_cdf(const float4* env, int W, int H, float* marg, float* cond, double* rowSums, float* total,
                    int* margGuide, int* condGuide, int* monotone, hipStream_t s) {
    hipLaunchKernelGGL(k_env_rows, dim3(H), dim3(kScanBlock), 0, s, env, W, H, cond, rowSums);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(kScanBlock), 0, s, rowSums, H, marg, total);
    launch_fill_i32(monotone, 1, 1, s);
    hipLaunchKernelGGL(k_env_guides, dim3(H + 1), dim3(256), 0, s, marg, cond, W, H, margGuide, condGuide, monotone);
}

// Footprint records (see envFoot): record (iyb + 1) * W + ix0 holds the RGB of
// texels (ix0, iy0), (ix1, iy0), (ix0, iy1), (ix1, iy1) in 12 floats and
// their pdf factors in 4, so that a lookup reads one aligned 64-byte record
// instead of two texel rows and four CDF entries.
__global__ void k_env_records(const float4* env, const float* marg, const float* cond, int W, int H, float total,
                              float4* rec) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)(H + 1) * W) return;
    const int ix0 = (int)(i % W), iyb = (int)(i / W) - 1;
    const int ix1 = (ix0 + 1 == W) ? 0 : ix0 + 1;
    const int iy0 = min(max(iyb, 0), H - 1), iy1 = min(max(iyb + 1, 0), H - 1);
    const float4 t00 = env[(size_t)iy0 * W + ix0], t10 = env[(size_t)iy0 * W + ix1];
    const float4 t01 = env[(size_t)iy1 * W + ix0], t11 = env[(size_t)iy1 * W + ix1];
    float4* r = rec + 4 * i;
    r[0] = make_float4(t00.x, t00.y, t00.z, t10.x);
    r[1] = make_float4(t10.y, t10.z, t01.x, t01.y);
    r[2] = make_float4(t01.z, t11.x, t11.y, t11.z);
    r[3] = make_float4(envPdfFactor(marg, cond, W, H, total, ix0, iy0), envPdfFactor(marg, cond, W, H, total, ix1, iy0),
                       envPdfFactor(marg, cond, W, H, total, ix0, iy1), envPdfFactor(marg, cond, W, H, total, ix1, iy1));
}

size_t env_record_bytes(int W, int H) { return (size_t)(H + 1) * (size_t)W * 4 * sizeof(float4); }

void launch_env_records(const float4* env, const float* marg, const float* cond, int W, int H, float total,
                        float4* rec, hipStream_t s) {
    const long long n = (long long)(H + 1) * W;
    hipLaunchKernelGGL(k_env_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, env, marg, cond, W, H, total,
                       rec);
}

// Resident blocks per CU of one category's shading kernel.
template <int T, bool B>
static int shadeOccupancy() {
    int a = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_shade<T, B>, kBlock, 0);
    return std::max(1, a);
}
template <int T>
static int shadeCamOccupancy() {
    int a = 0;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_shade<T, false, true>, kBlock, 0);
    return std::max(1, a);
}

// Resident blocks per CU of the traversal passes for one stack size and BVH width.
template <int L, int W>
static int travOccupancy() {
    int a = 0, b = 0;
    if constexpr (W == 2) {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend<false, L, 2>, kBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_connect<false, L, 2>, kBlock, 0);
    } else {
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_extend_pool<false, L, false, W>, kBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_connect_pool<false, L, false, W>, kBlock, 0);
    }
    return std::max(1, std::min(a, b));
}

template <int L, int W>
static int tailOccupancy() {
    int a = 0;
    if constexpr (W >= 4) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_tail_steps<L, W, false>, kBlock, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_tail<L, 2, false>, kBlock, 0);
    return std::max(1, a);
}

LaunchShape query_launch_shape(int device, int ldsStack, int traversalBlocksPerCU, int grabRows, int bvhWidth) {
    static int cachedDev = -1;
    static int occPB = 1;   // packet pass of batches
    static int cus = 1, occS = 1, occP = 1, occMT = 1, occL[3][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}};   // [width][stack]
    static int occT[3][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}};                                  // tail [width][stack]
    static int occCat[2][kNumCats] = {{1, 1, 1, 1, 1, 1}, {1, 1, 1, 1, 1, 1}};                  // shading [batch][category]
    static int occCam[kNumCats] = {1, 1, 1, 1, 1, 1};                                          // bounce-0 camera paths
    if (device != cachedDev) {
        hipDeviceProp_t prop;
        (void)hipGetDeviceProperties(&prop, device);
        cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
        int o = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_shade<BTE_BSDF_DISNEY>, kBlock, 0);
        occS = std::max(1, o);
        int o2 = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_extend_packet<false>, kBlock, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, k_extend_packet<true>, kBlock, 0);
        occP = std::max(1, o);
        occPB = std::max(1, o2);
        {   // the categoriser shades the misses
            int a = 0, b = 0;
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_categorize<false>, kBlock, 0);
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_categorize<true>, kBlock, 0);
            occCat[0][kCatMiss] = std::max(1, a);
            occCat[1][kCatMiss] = std::max(1, b);
        }
        {
            int a = 0;
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_categorize<false, true>, kBlock, 0);
            occCam[kCatMiss] = std::max(1, a);
        }
        occCam[BTE_BSDF_DIFFUSE] = shadeCamOccupancy<BTE_BSDF_DIFFUSE>();
        occCam[BTE_BSDF_DIELECTRIC] = shadeCamOccupancy<BTE_BSDF_DIELECTRIC>();
        occCam[BTE_BSDF_MIRROR] = shadeCamOccupancy<BTE_BSDF_MIRROR>();
        occCam[BTE_BSDF_MICROFACET] = shadeCamOccupancy<BTE_BSDF_MICROFACET>();
        occCam[BTE_BSDF_DISNEY] = shadeCamOccupancy<BTE_BSDF_DISNEY>();
        occCat[0][BTE_BSDF_DIFFUSE] = shadeOccupancy<BTE_BSDF_DIFFUSE, false>();
        occCat[0][BTE_BSDF_DIELECTRIC] = shadeOccupancy<BTE_BSDF_DIELECTRIC, false>();
        occCat[0][BTE_BSDF_MIRROR] = shadeOccupancy<BTE_BSDF_MIRROR, false>();
        occCat[0][BTE_BSDF_MICROFACET] = shadeOccupancy<BTE_BSDF_MICROFACET, false>();
        occCat[0][BTE_BSDF_DISNEY] = shadeOccupancy<BTE_BSDF_DISNEY, false>();
        occCat[1][BTE_BSDF_DIFFUSE] = shadeOccupancy<BTE_BSDF_DIFFUSE, true>();
        occCat[1][BTE_BSDF_DIELECTRIC] = shadeOccupancy<BTE_BSDF_DIELECTRIC, true>();
        occCat[1][BTE_BSDF_MIRROR] = shadeOccupancy<BTE_BSDF_MIRROR, true>();
        occCat[1][BTE_BSDF_MICROFACET] = shadeOccupancy<BTE_BSDF_MICROFACET, true>();
        occCat[1][BTE_BSDF_DISNEY] = shadeOccupancy<BTE_BSDF_DISNEY, true>();
        occL[0][0] = travOccupancy<8, 2>();
        occL[0][1] = travOccupancy<16, 2>();
        occL[0][2] = travOccupancy<32, 2>();
        occL[1][0] = travOccupancy<8, 4>();
        occL[1][1] = travOccupancy<16, 4>();
        occL[1][2] = travOccupancy<32, 4>();
        occL[2][0] = travOccupancy<8, 8>();
        occL[2][1] = travOccupancy<16, 8>();
        occL[2][2] = travOccupancy<32, 8>();
        occT[0][0] = tailOccupancy<8, 2>();
        occT[0][1] = tailOccupancy<16, 2>();
        occT[0][2] = tailOccupancy<32, 2>();
        occT[1][0] = tailOccupancy<8, 4>();
        occT[1][1] = tailOccupancy<16, 4>();
        occT[1][2] = tailOccupancy<32, 4>();
        occT[2][0] = tailOccupancy<8, 8>();
        occT[2][1] = tailOccupancy<16, 8>();
        occT[2][2] = tailOccupancy<32, 8>();
        cachedDev = device;
    }
    LaunchShape ls{};
    ls.ldsStack = (ldsStack == 8 || ldsStack == 16) ? ldsStack : 32;
    ls.bvhWidth = (bvhWidth == 4 || bvhWidth == 8) ? bvhWidth : 2;
    int occ = occL[ls.bvhWidth == 2 ? 0 : (ls.bvhWidth == 4 ? 1 : 2)][ls.ldsStack == 8 ? 0 : (ls.ldsStack == 16 ? 1 : 2)];
    if (traversalBlocksPerCU > 0) occ = std::min(occ, traversalBlocksPerCU);
    ls.grabRows = (grabRows == 1 || grabRows == 2 || grabRows == 8) ? grabRows : 4;
    ls.extendBlocks = cus * occ;
    ls.connectBlocks = cus * occ;
    ls.shadeBlocks = cus * occS;
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < kNumCats; ++k) ls.shadeCatBlocks[b][k] = cus * occCat[b][k];
    for (int k = 0; k < kNumCats; ++k) ls.shadeCamBlocks[k] = cus * occCam[k];
    ls.auxBlocks = cus * occMT;
    ls.packetBlocks = cus * occP;
    ls.packetBlocksBatch = cus * occPB;
    ls.tailBlocks = cus * occT[ls.bvhWidth == 2 ? 0 : (ls.bvhWidth == 4 ? 1 : 2)][ls.ldsStack == 8 ? 0 : (ls.ldsStack == 16 ? 1 : 2)];
    return ls;
}

int persistent_threads(const LaunchShape& ls) {
    return std::max(std::max(std::max(ls.extendBlocks, ls.connectBlocks), ls.auxBlocks), ls.tailBlocks) * kBlock;
}

int overflow_stack_ints(const LaunchShape& ls) { return persistent_threads(ls) * kOvfStack + 64; }

long long overflow_stack_ints_max(int device) {
    // every traversal width and LDS stack size the configuration can select,
    // at the largest grid any of them launches
    int most = 0;
    for (int w : {2, 4, 8})
        for (int st : {8, 16, 32}) most = std::max(most, persistent_threads(query_launch_shape(device, st, 0, 4, w)));
    return (long long)most * kOvfStack + 64;
}

}  // namespace bte
