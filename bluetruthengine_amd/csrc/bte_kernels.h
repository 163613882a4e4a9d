// bte_kernels.h — host-visible interface of the HIP kernels (bte_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "bte_types.h"

namespace bte {

// Read-only scene view passed by value to every kernel.
struct DevScene {
    const float4* nodes;          // BVH2, 4 float4 per node
    const float4* nodes4;         // BVH4, 8 float4 per node (null when not built)
    const float4* nodes8;         // BVH8, 16 float4 per node (null when not built)
    const float4* leafTris;       // 3 float4 per leaf-ordered triangle (v0.w = triangle id bits)
    const BteTriangle* tris;      // original order, 96 B each
    const int* triBsdf;
    const float4* bsdfs;          // 4 float4 per BSDF, type in p3.w bits
    const float* triEmission;     // 3 floats per triangle, may be null
    const uint8_t* triFlags;      // may be null
    const int* emitterOfTri;      // first emitter listing the triangle, -1 if none; may be null
    const BteEmitter* emitters;
    const int* emTriIdx;
    const float* emTriCdf;
    const float* sceneCdf;
    const BteSpotlight* spots;
    const float4* env;            // envW*envH texels (RGBA)
    const float* envMarg;
    const float* envCond;
    const int* envMargGuide;      // guide tables of the CDF inversion, null = plain binary search
    const int* envCondGuide;
    const float4* envRec;         // (envH + 1) * envW footprint records of 4 float4 (launch_env_records)
    const int* triMat;            // material id per triangle (texture index); null without textures
    const uint32_t* texels;       // RGBA8 texels of every albedo texture, back to back
    const int4* texInfo;          // [texCount]: texel offset, width, height (width 0 = no texture)
    const uint32_t* mrTexels;     // metallic-roughness maps, same layout
    const int4* mrInfo;           // [mrCount]
    int texCount, mrCount;
    BteCamera cam;
    int emitterCount;
    int spotCount;
    int envW, envH;
    float envIntensity, envRot, envTotal;
    int hasEnv;
    int width, height;
    int triCount;
};

// Counter block layout: kCtrStride ints per bounce iteration.
constexpr int kCtrRays = 0;       // rays entering this bounce
constexpr int kCtrExtendWork = 1;
constexpr int kCtrHits = 3;
constexpr int kCtrShadow = 4;                       // + kind: shadow rays of kind
constexpr int kCtrConnectWork = 4 + kMaxShadowKinds; // + kind
constexpr int kNumCats = 6;                          // shading categories: miss + the 5 BsdfTypes
constexpr int kCatMiss = 0;
constexpr int kCtrCat = 4 + 2 * kMaxShadowKinds;     // + category: rays queued for that category
constexpr int kCtrShadeWork = kCtrCat + kNumCats;    // + category
constexpr int kCtrSpotRays = kCtrShadeWork + kNumCats;  // spotlight shadow rays traced by the spot pass
constexpr int kCtrStride = kCtrSpotRays + 1;

// Per-scene queue counts of a batch, per bounce: the next bounce's rays, the
// category lists, the shadow records of each light class.
constexpr int kSegRays = 0;
constexpr int kSegCat = 1;                            // + category
constexpr int kSegShadow = kSegCat + kNumCats;        // + light class
constexpr int kSegQueues = kSegShadow + kMaxShadowKinds;
// A batch queue is walked in compact rows of 64 entries, each scene's rows
// rounded up to kSegRowAlign so that neither a work chunk (at most 8 rows) nor
// a categoriser block (16 rows) spans two scenes.
constexpr int kSegRowAlign = 16;

// Work counters of the persistent stages, per bounce: one group per stage
// (extend, shade per category, connect per light class) of kWorkXcds
// counters, each on its own 64-B line.  A block draws from counter
// blockIdx.x % kWorkXcds: workgroups are dealt to the 8 XCDs in turn, so each
// counter is hit by one XCD's waves and device-wide atomics spread over 8
// addresses instead of serialising on one.
constexpr int kWorkXcds = 8;
constexpr int kWorkPad = 16;
constexpr int kWorkExtend = 0;
constexpr int kWorkShade = 1;                        // + category
constexpr int kWorkConnect = kWorkShade + kNumCats;  // + light class
constexpr int kWorkTail = kWorkConnect + kMaxShadowKinds;
constexpr int kWorkGroups = kWorkTail + 1;
constexpr int kWorkStride = kWorkGroups * kWorkXcds * kWorkPad;   // ints per bounce

// Multi-scene batch: where each scene's paths and pixels live.  Path ids and
// queue slots share one layout: scene by scene, each scene's range starting at
// a multiple of kBatchChunk, so the scene of a slot is one byte-table lookup
// and a 64-slot row never holds two scenes.  Every queue (rays, category
// lists, shadow records) keeps each scene's entries at the start of that
// scene's range, with a count per (queue, scene).
constexpr int kBatchChunkShift = 12;
constexpr long long kBatchChunk = 1ll << kBatchChunkShift;
struct BatchSlot {
    long long pathBase;   // first path id of the scene (multiple of kBatchChunk)
    long long slotBase;   // first slot of the scene in a densely packed queue (unused by the kernels)
    int paths;            // pixels * frames of this pass
    int pixels;
    int pixelBase;        // first pixel of the scene in the batch's output buffers
    int pad;
};

// Wavefront buffers.  Path state lives in queue order and is compacted into
// the other half of each double buffer by the shade kernel.
struct WaveBuffers {
    float4* rayO[2];      // origin xyz, w = brdfPdf of the last scatter
    float4* rayD[2];      // direction xyz, w = eta
    // 2 float4 per slot, one aligned 32-byte record: (radiance xyz, rng state
    // bits), (throughput xyz, path id | flag bits)
    float4* aux[2];
    float4* hit;          // (triangle id bits, t, u, v) per queue slot
    int* catQ;            // [kNumCats * cap]: queue slots grouped by shading category
    uint32_t* cat;        // [cap]: shading category (bits 0-2) | BSDF table index << 3, per queue slot
    float4* finalRad;     // per path id
    float4* shO;          // [kinds * cap]: origin xyz, w = tMax
    float4* shD;          // [kinds * cap]: direction xyz, w = destination bits
    float4* shC;          // [kinds * cap]: contribution xyz
    // The spotlight kind holds one request per path instead: shO = (hit point,
    // destination bits), shD = (shading normal, triangle bits), shC =
    // (incoming direction, u), shX = (throughput, v).
    float4* shX;          // [cap], null without spotlights
    int* counters;        // kCtrStride * (maxDepth + 1)
    int* work;            // kWorkStride * (maxDepth + 1), zero at the start of a pass
    int* hitCapture;      // optional [P * maxRec]
    int* ovfStack;        // traversal stack overflow, per persistent thread
    unsigned long long* travCtr;  // optional: extend nodes, extend tris, connect nodes, connect tris
    long long cap;        // slots per buffer (paths in flight)
    int maxRec;
    int kinds;            // shadow kinds in use
    int kindArea, kindSpot, kindEnv;    // kind index of each light class (-1 absent)
    int catMask;          // bit t set: the scene has BSDF type t (its shading kernel runs)
    int grabRows;         // 64-slot rows a persistent wave takes per work-counter atomic (1, 2, 4, 8)
    float4* aovAlb;       // optional, per path id: albedo at the first non-delta hit (0 if none)
    float4* aovNrm;       // optional, per path id: shading normal there
    float4* aovAccAlb;    // per pixel running means of the two AOVs (with aovAlb)
    float4* aovAccNrm;
    const int* frameBase; // device int: frame index of the batch's first frame (set before each batch)
    // Single scenes: camPaths > 0 means the bounce-0 path state of the pass's
    // camPaths slots over camPixels pixels is not stored; the bounce-0
    // categoriser and shading kernels (their CAM instantiations) compute each
    // slot's camera path (cameraPath) where they would load it, and with
    // camRays set the bounce-0 rays are not stored either (the packet pass
    // computes them).
    long long camPaths;
    int camPixels;
    int camRays;
    // multi-scene batches only (null / 0 for a single scene)
    const uint8_t* sceneOfChunk;   // scene of each kBatchChunk-aligned block of slots / path ids
    const BatchSlot* slots;        // [sceneCount]
    int* segCount;                 // [(maxDepth + 1) * kSegQueues * sceneCount]
    int* segRows;                  // [(maxDepth + 1) * kSegQueues * (sceneCount + 1)]: first compact row of
                                   // each scene (its rows rounded up to kSegRowAlign), total at the end
    int sceneCount;
    int paddedSlots;               // end of the last scene's range
};

struct LaunchShape {
    int extendBlocks, shadeBlocks, connectBlocks, auxBlocks;
    int shadeCatBlocks[2][kNumCats];
    int shadeCamBlocks[kNumCats];       // grids of the bounce-0 (camera path) categoriser and shading kernels   // resident grid of each category's shading kernel [batch][category]
    int tailBlocks;       // resident grid of the tail pass (launch_tail)
    int packetBlocks;     // resident grid of the camera-ray packet pass (0: per-ray traversal on bounce 0)
    int packetBlocksBatch;               // the packet pass of multi-scene batches (its own occupancy)
    int ldsStack;         // traversal LDS stack entries per lane: 8, 16 or 32
    int grabRows;         // work chunk of the persistent kernels, in 64-slot rows
    int bvhWidth;         // 2 or 4: BVH traversed by extend / connect
};

// Launchers (all asynchronous on `stream`).
void launch_fill_i32(int* p, long long n, int value, hipStream_t s);
void launch_generate(const DevScene& sc, const WaveBuffers& wb, long long paths, int pixels, hipStream_t s);
// ev (optional): 4 events recorded before extend, after extend, after shade,
// after the connect passes.  scenes (device array, wb.sceneCount entries) selects
// the multi-scene kernels; sc is then unused.
void launch_bounce(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                   hipStream_t s, hipEvent_t* ev, const DevScene* scenes = nullptr);
void launch_present(const DevScene& sc, const WaveBuffers& wb, int pixels, int frames, float4* accum,
                    double* sum64, uint32_t* fb, hipStream_t s);
// Tail pass: finishes every path queued for `bounce` (all its remaining
// bounces with their shadow rays) in one persistent kernel, in place of
// launch_bounce for bounce..maxDepth-1.  Its per-bounce ray, hit and shadow
// counts land in the same counter blocks as launch_bounce's; ev (optional) as
// for launch_bounce, the whole pass timed as extension.
void launch_tail(const DevScene& sc, const WaveBuffers& wb, int bounce, int maxDepth, const LaunchShape& ls,
                 hipStream_t s, hipEvent_t* ev, const DevScene* scenes = nullptr);
// Multi-scene pass: paths of every scene (wb.slots) into its own slot range;
// present writes each scene's pixels at its pixelBase.
void launch_generate_batch(const DevScene* scenes, const WaveBuffers& wb, int maxScenePaths, long long totalSlots,
                           hipStream_t s);
// Row starts of queues [q, q + n) of bounce b from their per-scene counts; the
// ray and shadow queues' totals also go to the plain counters the host reads.
void launch_segment_rows(const WaveBuffers& wb, int bounce, int q, int n, hipStream_t s);
void launch_present_batch(const WaveBuffers& wb, int maxScenePixels, int frames, float4* accum, double* sum64,
                          uint32_t* fb, hipStream_t s);
void launch_megakernel(const DevScene& sc, int spp, int maxDepth, int frameIndex, float4* accum, double* sum64,
                       uint32_t* fb, int* ovfStack, hipStream_t s);
void launch_trace_rays(const DevScene& sc, const float* rays, const float* tmaxPerRay, long long n, float tmin,
                       float tmax, int anyHit, int* tri, float* t, float* u, float* v, int* ovfStack,
                       hipStream_t s);
// Allocation-free form of launch_trace_rays: scratchPtr is the device
// address of the library's static scratch array (noalloc_scratch), whose
// first int is cleared here; the BVH2 must be at most noalloc_max_bvh_depth()
// levels deep.
void launch_trace_rays_noalloc(const DevScene& sc, const float* rays, const float* tmaxPerRay, long long n,
                               float tmin, float tmax, int anyHit, int* tri, float* t, float* u, float* v,
                               int* scratchPtr, hipStream_t s);
// Device address of the static scratch array on the current device (looked up
// once per device with hipGetSymbolAddress).
hipError_t noalloc_scratch(int** ptr);
size_t noalloc_scratch_bytes();
int noalloc_max_bvh_depth();
void launch_env_cdf(const float4* env, int W, int H, float* marg, float* cond, double* rowSums, float* total,
                    int* margGuide, int* condGuide, int* monotone, hipStream_t s);
// Footprint records of the bilinear environment lookup and its pdf, built
// from the texels and the CDFs of launch_env_cdf (total: their power sum).
size_t env_record_bytes(int W, int H);
void launch_env_records(const float4* env, const float* marg, const float* cond, int W, int H, float total,
                        float4* rec, hipStream_t s);
// traversalBlocksPerCU 0 = as many as occupancy allows; grabRows 1, 2, 4 or 8 (else 4).
LaunchShape query_launch_shape(int device, int ldsStack, int traversalBlocksPerCU, int grabRows, int bvhWidth);
int persistent_threads(const LaunchShape& ls);
int overflow_stack_ints(const LaunchShape& ls);   // overflow-stack ints a launch of shape ls indexes
// Overflow-stack ints that cover every launch shape query_launch_shape can
// return on `device` (all BVH widths and LDS stack sizes, full occupancy).
long long overflow_stack_ints_max(int device);

}  // namespace bte
