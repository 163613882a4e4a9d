// bte_api.cpp — C ABI (include/bte_api.h) over the HIP kernels.
//
// Reference counterparts: include/render_kernel.h:96-211 (API),
// src/host/scene_upload.cu (uploads), src/host/optix_setup.cu:215-292 (BVH
// build), src/host/wavefront_host.cu:155-236 (bounce loop),
// src/render_kernel.cu:33-183 (frame dispatch, reset, cleanup).
//
// Differences by design: status codes instead of exit(); the bounce loop keeps
// every count on the device (no per-bounce readback); several frames are traced
// concurrently as one wavefront pass; NEE contributions are applied by
// deterministic per-light-class connect passes instead of float atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bte_bvh.h"
#include "bte_bvh_gpu.h"
#include "bte_internal.h"
#include "bte_kernels.h"

namespace bte {

char g_err[1024] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

namespace {

// Everything bte_init_* sets up for one scene.  The context holds the current
// scene; bte_batch_add_scene moves it into a batch slot.
struct SceneData {
    // host copies (the BVH is built from triangles on the host)
    std::vector<BteTriangle> tris;
    std::vector<BteFloat3> materials;
    std::vector<int> triMaterialIds;
    std::vector<BteBsdf> bsdfsHost;
    std::vector<int> triBsdfHost;
    std::vector<BteEmitter> emittersHost;
    std::vector<int> emTriIdxHost;
    int emitterCount = 0, emTriIdxCount = 0, emTriCdfCount = 0;
    int spotCount = 0;
    bool haveCamera = false;
    BteCamera camera{};
    int sceneW = 0, sceneH = 0;

    // device scene
    DBuf dTris, dMaterials, dTriMat, dTriBsdf, dBsdfs, dEmission, dFlags, dEmitterOfTri;
    DBuf dTexels, dTexInfo, dMrTexels, dMrInfo;   // texture maps (rt_material.cuh:26-56)
    int texCount = 0, mrCount = 0;
    DBuf dEmitters, dEmTriIdx, dEmTriCdf, dSceneCdf, dSpots;
    DBuf dNodes, dNodes4, dNodes8, dLeafTris;
    int bvh4Depth = 0, bvh8Depth = 0;
    DBuf dEnv, dEnvMarg, dEnvCond, dEnvRowSums, dEnvTotal, dEnvMargGuide, dEnvCondGuide, dEnvMono;
    DBuf dEnvRec;   // footprint records of the bilinear lookup and its pdf (launch_env_records)
    bool envGuides = false;
    bool haveEmission = false, haveFlags = false;
    bool bvhDirty = true;
    int bvhNodes = 0, bvhDepth = 0;
    BteAccelStats accel{};
    int envW = 0, envH = 0;
    float envIntensity = 0.0f, envRot = 0.0f, envTotal = 0.0f;
    bool hasEnv = false;
};

// Multi-scene batch: its scenes, their output buffers and its frame counter.
struct BatchScene {
    std::unique_ptr<SceneData> data;
    int w = 0, h = 0;
    long long pixelBase = 0;
};

// Library defaults (bte_get_default_config): reference depth and megakernel
// paths per pixel, automatic frames in flight, 16-entry LDS stacks, graphs on,
// 4-row work chunks, BVH4, tail pass from bounce 6 (measured on the 1080p
// env-lit and 720p terrain workloads).
constexpr BteConfig kDefaultConfig{200, 5, 0, 32768, 16, 1, 4, 0, 4, 6};

struct Context : SceneData {
    bool deviceReady = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::vector<BatchScene> batch;
    unsigned long long batchGen = 0;          // changes whenever the slot list changes
    long long batchPixels = 0;
    uint32_t batchFrame = 0;
    DBuf dBatchScenes, dBatchSlots, dBatchChunk, dBatchAccum, dBatchFb, dBatchSum64, dBatchAovAlb, dBatchAovNrm,
        dBatchDenoised, dBatchSegCount, dBatchSegRows;
    long long batchLayoutKey = -1;            // frames per pass the slot table was built for
    int batchLastFrames = 0;                  // frames of the batch's last wavefront pass
    unsigned long long batchLayoutGen = ~0ull;
    bool batchDenoised = false;

    // output
    DBuf dAccum, dFb, dSum64;
    uint32_t* externalFb = nullptr;
    int accW = 0, accH = 0;
    uint32_t frameIndex = 0;
    bool accum64 = false;
    int mode = BTE_MODE_WAVEFRONT;

    // wavefront buffers
    DBuf dRayO[2], dRayD[2], dAux[2], dHit, dCatQ, dCat, dFinal, dShO, dShD, dShC, dShX, dCounters,
        dHitCap, dOvf;
    DBuf dFrameBase, dWork;
    bool aov = false;                 // albedo / normal AOVs of the first non-delta hit
    DBuf dAovAlb, dAovNrm, dAovAccAlb, dAovAccNrm;

    // captured bounce-loop segments, keyed by every value their launches bake in
    struct GraphEntry {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
        unsigned long long lastUse = 0;
    };
    std::vector<GraphEntry> graphs;
    unsigned long long graphTick = 0;
    long long graphLaunches = 0;
    long long cap = 0;
    int capKinds = 0;
    int hitCaptureRec = 0;
    long long lastPassPaths = 0;
    int lastPassFrames = 0;

    BteConfig cfg = kDefaultConfig;
    int accelBuilder = 1;                      // 0: host binned SAH, 1: GPU clustering
    BteAccelStats lastAccel{};                 // of the last BVH build (any scene)
    BteStats stats{};
    int instr = 0;
    DBuf dTrav;
    std::vector<hipEvent_t> events;
};

Context& ctx() {
    static Context c;
    return c;
}

int ensureDevice() {
    Context& c = ctx();
    if (c.deviceReady) return BTE_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(BTE_ERR_DEVICE, "no HIP device available");
    }
    HIPCHK(hipGetDevice(&c.device));
    HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c.ev0));
    HIPCHK(hipEventCreate(&c.ev1));
    c.deviceReady = true;
    return BTE_OK;
}

#define ENSURE_DEVICE()              \
    do {                             \
        int rc_ = ensureDevice();    \
        if (rc_) return rc_;         \
    } while (0)

int rebuildEmitterOfTri() {
    SceneData& c = ctx();
    const int T = (int)c.tris.size();
    if (c.emitterCount <= 0 || T == 0) {
        c.dEmitterOfTri.release();
        return BTE_OK;
    }
    std::vector<int> eot((size_t)T, -1);
    for (int e = 0; e < c.emitterCount; ++e) {
        const BteEmitter& E = c.emittersHost[e];
        for (int j = 0; j < E.triCount; ++j) {
            const int t = c.emTriIdxHost[E.triIndexOffset + j];
            if (t >= 0 && t < T && eot[t] < 0) eot[t] = e;
        }
    }
    return c.dEmitterOfTri.upload(eot.data(), eot.size() * sizeof(int));
}

int ensureAccel(SceneData& c) {
    if (!c.bvhDirty) return BTE_OK;
    const auto t0 = std::chrono::steady_clock::now();
    const int T = (int)c.tris.size();
    if (ctx().accelBuilder == 1 && T >= 2) {
        GpuBvh g;
        CHECK_RC(build_bvh_gpu(c.dTris.as<BteTriangle>(), T, ctx().stream, g));
        // a tree too deep for the traversal stacks is built again on the host
        if (g.depth2 <= kMaxBvhDepth && g.depth4 <= g.depth2) {
            c.dNodes = std::move(g.dNodes2);
            c.dNodes4 = std::move(g.dNodes4);
            c.bvh4Depth = g.depth4;
            c.bvh8Depth = g.depth8;
            if (7 * g.depth8 + 1 <= kStackSize) c.dNodes8 = std::move(g.dNodes8);
            else c.dNodes8.release();
            c.dLeafTris = std::move(g.dLeafTris);
            c.bvhNodes = g.nodes2;
            c.bvhDepth = g.depth2;
            c.bvhDirty = false;
            c.accel = BteAccelStats{std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                                        .count(),
                                    1, T, g.nodes2, g.nodes4, c.dNodes8.p ? g.nodes8 : 0, g.depth2, g.depth4,
                                    g.depth8};
            ctx().lastAccel = c.accel;
            return BTE_OK;
        }
    }
    BvhBuild b = build_bvh2(c.tris.data(), T);
    CHECK_RC(c.dNodes.upload(b.nodes.data(), b.nodes.size() * sizeof(BvhHostNode)));
    {
        const std::vector<Bvh4HostNode> n4 = collapse_bvh4(b.nodes, c.bvh4Depth);
        CHECK_RC(c.dNodes4.upload(n4.data(), n4.size() * sizeof(Bvh4HostNode)));
        c.accel.nodes4 = (int)n4.size();
    }
    {
        // the BVH8 needs up to 7 stack entries per level; a tree too deep for
        // the traversal stack is not kept (width 8 then runs on the BVH4)
        const std::vector<Bvh8HostNode> n8 = collapse_bvh8(b.nodes, c.bvh8Depth);
        if (7 * c.bvh8Depth + 1 <= kStackSize) CHECK_RC(c.dNodes8.upload(n8.data(), n8.size() * sizeof(Bvh8HostNode)));
        else c.dNodes8.release();
        c.accel.nodes8 = c.dNodes8.p ? (int)n8.size() : 0;
    }
    std::vector<float> lt = std::move(b.leafTris);
    if (lt.empty()) lt.assign(12, 0.0f);
    CHECK_RC(c.dLeafTris.upload(lt.data(), lt.size() * sizeof(float)));
    c.bvhNodes = (int)b.nodes.size();
    c.bvhDepth = b.depth;
    c.bvhDirty = false;
    c.accel.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    c.accel.builder = 0;
    c.accel.triangles = T;
    c.accel.nodes2 = c.bvhNodes;
    c.accel.depth2 = c.bvhDepth;
    c.accel.depth4 = c.bvh4Depth;
    c.accel.depth8 = c.bvh8Depth;
    ctx().lastAccel = c.accel;
    return BTE_OK;
}

DevScene makeScene(const SceneData& c, int width, int height) {
    DevScene s;
    std::memset(&s, 0, sizeof s);   // padding too: scenes are compared bytewise as graph keys
    s.nodes = c.dNodes.as<float4>();
    s.nodes4 = c.dNodes4.as<float4>();
    s.nodes8 = c.dNodes8.as<float4>();
    s.leafTris = c.dLeafTris.as<float4>();
    s.tris = c.dTris.as<BteTriangle>();
    s.triBsdf = c.dTriBsdf.as<int>();
    s.bsdfs = c.dBsdfs.as<float4>();
    s.triEmission = c.haveEmission ? c.dEmission.as<float>() : nullptr;
    s.triFlags = c.haveFlags ? c.dFlags.as<uint8_t>() : nullptr;
    s.emitterOfTri = c.emitterCount > 0 ? c.dEmitterOfTri.as<int>() : nullptr;
    s.emitters = c.dEmitters.as<BteEmitter>();
    s.emTriIdx = c.dEmTriIdx.as<int>();
    s.emTriCdf = c.dEmTriCdf.as<float>();
    s.sceneCdf = c.dSceneCdf.as<float>();
    s.spots = c.dSpots.as<BteSpotlight>();
    s.env = c.dEnv.as<float4>();
    s.envMarg = c.dEnvMarg.as<float>();
    s.envCond = c.dEnvCond.as<float>();
    s.envMargGuide = c.envGuides ? c.dEnvMargGuide.as<int>() : nullptr;
    s.envCondGuide = c.envGuides ? c.dEnvCondGuide.as<int>() : nullptr;
    s.envRec = c.dEnvRec.as<float4>();
    if (c.texCount > 0) {
        s.triMat = c.dTriMat.as<int>();
        s.texels = c.dTexels.as<uint32_t>();
        s.texInfo = c.dTexInfo.as<int4>();
        s.texCount = c.texCount;
        s.mrTexels = c.mrCount > 0 ? c.dMrTexels.as<uint32_t>() : nullptr;
        s.mrInfo = c.mrCount > 0 ? c.dMrInfo.as<int4>() : nullptr;
        s.mrCount = c.mrCount;
    }
    s.cam = c.camera;
    s.emitterCount = c.emitterCount;
    s.spotCount = c.spotCount;
    s.envW = c.envW;
    s.envH = c.envH;
    s.envIntensity = c.envIntensity;
    s.envRot = c.envRot;
    s.envTotal = c.envTotal;
    s.hasEnv = c.hasEnv ? 1 : 0;
    s.width = width;
    s.height = height;
    s.triCount = (int)c.tris.size();
    return s;
}

int validateForRender(const SceneData& c) {
    if (c.tris.empty()) return fail(BTE_ERR_STATE, "no triangles uploaded (bte_init_scene)");
    if (!c.haveCamera) return fail(BTE_ERR_STATE, "no camera uploaded (bte_init_camera)");
    if (c.bsdfsHost.empty() || c.triBsdfHost.size() != c.tris.size())
        return fail(BTE_ERR_STATE, "BSDF table missing or triangle count mismatch (bte_init_bsdfs)");
    if (c.haveEmission && c.dEmission.bytes < c.tris.size() * 12)
        return fail(BTE_ERR_STATE, "triangle emission count does not match the triangles");
    if (c.haveFlags && c.dFlags.bytes < c.tris.size())
        return fail(BTE_ERR_STATE, "triangle emitter flags count does not match the triangles");
    if (c.emitterCount > 0) {
        for (int e = 0; e < c.emitterCount; ++e) {
            const BteEmitter& E = c.emittersHost[e];
            if (E.triCount < 0 || E.triIndexOffset < 0 || E.triIndexOffset + E.triCount > c.emTriIdxCount ||
                E.cdfOffset < 0 || E.cdfOffset + E.triCount + 1 > c.emTriCdfCount)
                return fail(BTE_ERR_INVALID, "emitter %d references data outside the emitter tables", e);
        }
        for (int i = 0; i < c.emTriIdxCount; ++i)
            if (c.emTriIdxHost[i] < 0 || c.emTriIdxHost[i] >= (int)c.tris.size())
                return fail(BTE_ERR_INVALID, "emitter triangle index %d out of range", c.emTriIdxHost[i]);
    }
    return BTE_OK;
}

int resetAccum(int w, int h) {
    Context& c = ctx();
    if (w <= 0 || h <= 0) return fail(BTE_ERR_INVALID, "invalid resolution %dx%d", w, h);
    const size_t N = (size_t)w * (size_t)h;
    CHECK_RC(c.dAccum.ensure(N * sizeof(float4)));
    CHECK_RC(c.dFb.ensure(N * sizeof(uint32_t)));
    HIPCHK(hipMemsetAsync(c.dAccum.p, 0, N * sizeof(float4), c.stream));
    if (c.aov) {
        CHECK_RC(c.dAovAccAlb.ensure(N * sizeof(float4)));
        CHECK_RC(c.dAovAccNrm.ensure(N * sizeof(float4)));
        HIPCHK(hipMemsetAsync(c.dAovAccAlb.p, 0, N * sizeof(float4), c.stream));
        HIPCHK(hipMemsetAsync(c.dAovAccNrm.p, 0, N * sizeof(float4), c.stream));
    }
    if (c.accum64) {
        CHECK_RC(c.dSum64.ensure(N * 3 * sizeof(double)));
        HIPCHK(hipMemsetAsync(c.dSum64.p, 0, N * 3 * sizeof(double), c.stream));
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    c.accW = w;
    c.accH = h;
    c.frameIndex = 0;
    return BTE_OK;
}

int kindsInUse(int& kindArea, int& kindSpot, int& kindEnv) {
    Context& c = ctx();
    int k = 0;
    kindArea = c.emitterCount > 0 ? k++ : -1;
    kindSpot = c.spotCount > 0 ? k++ : -1;
    kindEnv = c.hasEnv ? k++ : -1;
    return k;
}

size_t bytesPerPath(int kinds, bool spots) {
    return 2 * (2 * sizeof(float4) + 2 * sizeof(float4)) + 2 * sizeof(float4) + kNumCats * sizeof(int) + sizeof(uint32_t) +
           (size_t)kinds * 3 * sizeof(float4) + (spots ? sizeof(float4) : 0) + (ctx().aov ? 2 * sizeof(float4) : 0);
}

int ensureWaveBuffers(long long cap, int kinds, bool spots, int maxDepth, WaveBuffers& wb) {
    Context& c = ctx();
    const size_t P = (size_t)cap;
    for (int i = 0; i < 2; ++i) {
        CHECK_RC(c.dRayO[i].ensure(P * sizeof(float4)));
        CHECK_RC(c.dRayD[i].ensure(P * sizeof(float4)));
        CHECK_RC(c.dAux[i].ensure(2 * P * sizeof(float4)));
    }
    CHECK_RC(c.dHit.ensure(P * sizeof(float4)));
    CHECK_RC(c.dCatQ.ensure((size_t)kNumCats * P * sizeof(int)));
    CHECK_RC(c.dCat.ensure(P * sizeof(uint32_t)));
    CHECK_RC(c.dFinal.ensure(P * sizeof(float4)));
    const size_t K = (size_t)std::max(kinds, 1);
    CHECK_RC(c.dShO.ensure(K * P * sizeof(float4)));
    CHECK_RC(c.dShD.ensure(K * P * sizeof(float4)));
    CHECK_RC(c.dShC.ensure(K * P * sizeof(float4)));
    if (spots) CHECK_RC(c.dShX.ensure(P * sizeof(float4)));
    CHECK_RC(c.dCounters.ensure((size_t)kCtrStride * (size_t)(maxDepth + 1) * sizeof(int)));
    CHECK_RC(c.dWork.ensure((size_t)kWorkStride * (size_t)(maxDepth + 1) * sizeof(int)));
    CHECK_RC(c.dFrameBase.ensure(sizeof(int)));
    if (c.aov) {
        CHECK_RC(c.dAovAlb.ensure(P * sizeof(float4)));
        CHECK_RC(c.dAovNrm.ensure(P * sizeof(float4)));
    }
    CHECK_RC(c.dOvf.ensure((size_t)overflow_stack_ints_max(c.device) * sizeof(int)));
    if (c.hitCaptureRec > 0) CHECK_RC(c.dHitCap.ensure(P * (size_t)c.hitCaptureRec * sizeof(int)));
    c.cap = cap;
    c.capKinds = kinds;
    for (int i = 0; i < 2; ++i) {
        wb.rayO[i] = c.dRayO[i].as<float4>();
        wb.rayD[i] = c.dRayD[i].as<float4>();
        wb.aux[i] = c.dAux[i].as<float4>();
    }
    wb.hit = c.dHit.as<float4>();
    wb.catQ = c.dCatQ.as<int>();
    wb.cat = c.dCat.as<uint32_t>();
    wb.finalRad = c.dFinal.as<float4>();
    wb.shO = c.dShO.as<float4>();
    wb.shD = c.dShD.as<float4>();
    wb.shC = c.dShC.as<float4>();
    wb.shX = spots ? c.dShX.as<float4>() : nullptr;
    wb.counters = c.dCounters.as<int>();
    wb.work = c.dWork.as<int>();
    wb.frameBase = c.dFrameBase.as<int>();
    wb.aovAlb = c.aov ? c.dAovAlb.as<float4>() : nullptr;
    wb.aovNrm = c.aov ? c.dAovNrm.as<float4>() : nullptr;
    wb.aovAccAlb = c.aov ? c.dAovAccAlb.as<float4>() : nullptr;
    wb.aovAccNrm = c.aov ? c.dAovAccNrm.as<float4>() : nullptr;
    wb.hitCapture = c.hitCaptureRec > 0 ? c.dHitCap.as<int>() : nullptr;
    wb.ovfStack = c.dOvf.as<int>() + 64;
    wb.cap = cap;
    wb.maxRec = c.hitCaptureRec;
    return BTE_OK;
}

void dropGraphs() {
    Context& c = ctx();
    for (auto& g : c.graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    c.graphs.clear();
}

template <class T>
void appendKey(std::vector<unsigned char>& key, const T& v) {
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
    key.insert(key.end(), b, b + sizeof(T));
}

// Runs `emit` (a sequence of kernel launches on the context stream) either
// directly or as a captured HIP graph that is cached under `key` and replayed
// whenever the same key comes back.  Capture records kernel nodes only (fills
// are kernels, never memset nodes).
template <class F>
int runSegment(bool useGraph, const std::vector<unsigned char>& key, F&& emit) {
    Context& c = ctx();
    if (!useGraph) {
        emit();
        return BTE_OK;
    }
    for (auto& g : c.graphs) {
        if (g.key == key) {
            g.lastUse = ++c.graphTick;
            HIPCHK(hipGraphLaunch(g.exec, c.stream));
            ++c.graphLaunches;
            return BTE_OK;
        }
    }
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    emit();
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(c.stream, &graph);
    if (ce != hipSuccess || !graph) {
        (void)hipGetLastError();
        return fail(BTE_ERR_DEVICE, "stream capture of the bounce loop failed: %s", hipGetErrorString(ce));
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        (void)hipGetLastError();
        return fail(BTE_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ie));
    }
    constexpr size_t kMaxGraphs = 32;
    if (c.graphs.size() >= kMaxGraphs) {
        auto oldest = std::min_element(c.graphs.begin(), c.graphs.end(),
                                       [](const Context::GraphEntry& x, const Context::GraphEntry& y) {
                                           return x.lastUse < y.lastUse;
                                       });
        (void)hipGraphExecDestroy(oldest->exec);
        c.graphs.erase(oldest);
    }
    c.graphs.push_back(Context::GraphEntry{key, exec, ++c.graphTick});
    HIPCHK(hipGraphLaunch(exec, c.stream));
    ++c.graphLaunches;
    return BTE_OK;
}

// Bounce from which the tail pass finishes the paths (maxDepth: never).
int tailBounce(int maxDepth) {
    const int t = ctx().cfg.tailBounce;
    return (t > 0 && t < maxDepth) ? t : maxDepth;
}

int renderWavefront(int nframes) {
    Context& c = ctx();
    const int N = c.accW * c.accH;
    int kindArea, kindSpot, kindEnv;
    const int kinds = kindsInUse(kindArea, kindSpot, kindEnv);
    const int maxDepth = std::max(1, c.cfg.maxDepth);
    int F = c.cfg.framesInFlight;
    if (F <= 0) {
        const double budget = (double)(c.cfg.memoryBudgetMB > 0 ? c.cfg.memoryBudgetMB : 32768) * 1048576.0;
        F = (int)std::max(1.0, std::floor(budget / ((double)N * (double)bytesPerPath(kinds, kindSpot >= 0))));
    }
    F = std::max(1, std::min(F, nframes));
    // path ids occupy 30 bits of the path word
    if ((long long)N * F > (1ll << 30)) F = std::max(1, (int)((1ll << 30) / N));
    if ((long long)N > (1ll << 30)) return fail(BTE_ERR_INVALID, "more than 2^30 pixels per frame");
    WaveBuffers wb;
    std::memset(&wb, 0, sizeof wb);   // padding too: compared bytewise as graph keys
    CHECK_RC(ensureWaveBuffers((long long)N * F, kinds, kindSpot >= 0, maxDepth, wb));
    wb.kinds = kinds;
    wb.travCtr = nullptr;
    if (c.instr & BTE_INSTR_TRAVERSAL_COUNTS) {
        CHECK_RC(c.dTrav.ensure(4 * sizeof(unsigned long long)));
        HIPCHK(hipMemsetAsync(c.dTrav.p, 0, 4 * sizeof(unsigned long long), c.stream));
        wb.travCtr = c.dTrav.as<unsigned long long>();
    }
    const bool timed = (c.instr & BTE_INSTR_KERNEL_TIMES) != 0;
    if (timed && c.events.size() < (size_t)4 * maxDepth) {
        for (size_t i = c.events.size(); i < (size_t)4 * maxDepth; ++i) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c.events.push_back(e);
        }
    }
    wb.catMask = 0;
    for (const BteBsdf& b : c.bsdfsHost) wb.catMask |= 1 << b.type;
    wb.kindArea = kindArea;
    wb.kindSpot = kindSpot;
    wb.kindEnv = kindEnv;
    const DevScene sc = makeScene(c, c.accW, c.accH);
    LaunchShape ls = query_launch_shape(c.device, c.cfg.traversalStack > 0 ? c.cfg.traversalStack : 16,
                                        c.cfg.traversalBlocksPerCU, c.cfg.workChunk, c.cfg.bvhWidth);
    if (ls.bvhWidth == 8 && !c.dNodes8.p) ls.bvhWidth = 4;   // BVH8 too deep for the stack
    if ((size_t)overflow_stack_ints(ls) * sizeof(int) > c.dOvf.bytes)
        return fail(BTE_ERR_STATE, "traversal overflow stack smaller than the launch shape needs");
    wb.grabRows = ls.grabRows;
    // instrumented runs record events between launches: those stay uncaptured
    const bool useGraph = c.cfg.graphs != 0 && c.instr == 0;
    // traversal counts are kept by the per-bounce passes only
    const int tailAt = wb.travCtr ? maxDepth : tailBounce(maxDepth);
    uint32_t* fb = c.externalFb ? c.externalFb : c.dFb.as<uint32_t>();
    double* sum64 = c.accum64 ? c.dSum64.as<double>() : nullptr;
    const size_t ctrInts = (size_t)kCtrStride * (size_t)(maxDepth + 1);
    std::vector<int> ctrHost(ctrInts);

    c.stats = BteStats{};
    c.stats.framesInFlight = F;
    c.stats.bvhNodes = c.bvhNodes;
    c.stats.bvhDepth = c.bvhDepth;
    HIPCHK(hipEventRecord(c.ev0, c.stream));
    const long long graphLaunches0 = c.graphLaunches;
    int done = 0;
    while (done < nframes) {
        const int frames = std::min(F, nframes - done);
        const long long paths = (long long)N * frames;
        // the bounce-0 passes compute each slot's camera path instead of
        // loading it (k_generate stores none)
        wb.camPaths = paths;
        wb.camPixels = N;
        // rays too when the packet pass traces bounce 0 (it computes them)
        wb.camRays = (ls.bvhWidth == 4 && ls.packetBlocks > 0 && !wb.travCtr) ? 1 : 0;
        // the batch's first frame index lives in device memory, so captured
        // segments replay unchanged from frame to frame
        launch_fill_i32(c.dFrameBase.as<int>(), 1, (int)c.frameIndex, c.stream);
        // Bounces run in segments of kSegment; between segments the host reads
        // the next bounce's ray count and stops early once no path is left.
        // The segment holding the tail bounce ends with the tail pass.
        constexpr int kSegment = 16;
        int b = 0, bTimed = 0;
        bool presented = false;
        for (int b0 = 0; b0 < maxDepth; b0 += kSegment) {
            const int b1 = std::min(b0 + kSegment, maxDepth);
            const bool tail = tailAt >= b0 && tailAt < b1;
            const int bEnd = tail ? tailAt : b1;
            const bool last = b1 == maxDepth || tail;
            std::vector<unsigned char> key;
            if (useGraph) {
                appendKey(key, sc);
                appendKey(key, wb);
                appendKey(key, ls);
                const long long params[] = {paths, N, frames, b0, b1, maxDepth, tailAt, (long long)(uintptr_t)fb,
                                            (long long)(uintptr_t)sum64, (long long)(uintptr_t)c.dAccum.p};
                appendKey(key, params);
            }
            CHECK_RC(runSegment(useGraph, key, [&]() {
                if (b0 == 0) {
                    launch_fill_i32(wb.counters, (long long)ctrInts, 0, c.stream);
                    launch_fill_i32(wb.work, (long long)kWorkStride * (maxDepth + 1), 0, c.stream);
                    if (wb.hitCapture) launch_fill_i32(wb.hitCapture, paths * wb.maxRec, -2, c.stream);
                    launch_generate(sc, wb, paths, N, c.stream);
                }
                for (int bb = b0; bb < bEnd; ++bb)
                    launch_bounce(sc, wb, bb, maxDepth, ls, c.stream, timed ? &c.events[4 * (size_t)bb] : nullptr);
                if (tail)
                    launch_tail(sc, wb, tailAt, maxDepth, ls, c.stream, timed ? &c.events[4 * (size_t)tailAt] : nullptr);
                if (last) launch_present(sc, wb, N, frames, c.dAccum.as<float4>(), sum64, fb, c.stream);
            }));
            b = tail ? maxDepth : b1;
            bTimed = tail ? tailAt + 1 : b1;
            if (last) {
                presented = true;
                break;
            }
            int next = 0;
            HIPCHK(hipMemcpyAsync(&next, wb.counters + (size_t)b1 * kCtrStride + kCtrRays, sizeof(int),
                                  hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (next == 0) break;
        }
        if (!presented) launch_present(sc, wb, N, frames, c.dAccum.as<float4>(), sum64, fb, c.stream);
        HIPCHK(hipMemcpyAsync(ctrHost.data(), wb.counters, ctrInts * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        HIPCHK(hipGetLastError());
        for (int i = 0; i < b; ++i) {
            const int* cb = ctrHost.data() + (size_t)i * kCtrStride;
            c.stats.extensionRays += cb[kCtrRays];
            c.stats.hits += cb[kCtrHits];                                     // tail pass
            for (int k = 1; k < kNumCats; ++k) c.stats.hits += cb[kCtrCat + k];   // per-bounce passes
            for (int k = 0; k < kinds; ++k) c.stats.shadowRays += (k == kindSpot) ? cb[kCtrSpotRays] : cb[kCtrShadow + k];
            if (i < 64) c.stats.raysPerBounce[i] += cb[kCtrRays];
        }
        c.stats.bounces = std::max(c.stats.bounces, b);
        if (timed) {
            for (int i = 0; i < bTimed; ++i) {
                float e = 0.0f, sh = 0.0f, cn = 0.0f;
                HIPCHK(hipEventElapsedTime(&e, c.events[4 * i], c.events[4 * i + 1]));
                if (i == tailAt) {            // the tail pass: its own figure
                    c.stats.tailMs += e;
                    ++c.stats.tailLaunches;
                    continue;
                }
                HIPCHK(hipEventElapsedTime(&sh, c.events[4 * i + 1], c.events[4 * i + 2]));
                HIPCHK(hipEventElapsedTime(&cn, c.events[4 * i + 2], c.events[4 * i + 3]));
                c.stats.extendMs += e;
                c.stats.shadeMs += sh;
                c.stats.connectMs += cn;
                ++c.stats.extendLaunches;
                c.stats.connectLaunches += kinds;
            }
        }
        c.stats.pathsTraced += paths;
        c.frameIndex += (uint32_t)frames;
        c.lastPassPaths = paths;
        c.lastPassFrames = frames;
        done += frames;
    }
    HIPCHK(hipEventRecord(c.ev1, c.stream));
    HIPCHK(hipEventSynchronize(c.ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.stats.renderMs = ms;
    c.stats.graphLaunches = (int)(c.graphLaunches - graphLaunches0);
    if (wb.travCtr) {
        unsigned long long tc[4];
        HIPCHK(hipMemcpy(tc, c.dTrav.p, sizeof(tc), hipMemcpyDeviceToHost));
        c.stats.extendNodes = (long long)tc[0];
        c.stats.extendTris = (long long)tc[1];
        c.stats.connectNodes = (long long)tc[2];
        c.stats.connectTris = (long long)tc[3];
    }
    return BTE_OK;
}

int renderMegakernel(int nframes) {
    Context& c = ctx();
    const DevScene sc = makeScene(c, c.accW, c.accH);
    CHECK_RC(c.dOvf.ensure((size_t)overflow_stack_ints_max(c.device) * sizeof(int)));
    uint32_t* fb = c.externalFb ? c.externalFb : c.dFb.as<uint32_t>();
    double* sum64 = c.accum64 ? c.dSum64.as<double>() : nullptr;
    const int spp = std::max(1, c.cfg.megakernelSpp);
    c.stats = BteStats{};
    c.stats.bvhNodes = c.bvhNodes;
    c.stats.bvhDepth = c.bvhDepth;
    HIPCHK(hipEventRecord(c.ev0, c.stream));
    for (int f = 0; f < nframes; ++f) {
        launch_megakernel(sc, spp, std::max(1, c.cfg.maxDepth), (int)c.frameIndex, c.dAccum.as<float4>(), sum64, fb,
                          c.dOvf.as<int>(), c.stream);
        c.frameIndex++;
    }
    HIPCHK(hipEventRecord(c.ev1, c.stream));
    HIPCHK(hipEventSynchronize(c.ev1));
    HIPCHK(hipGetLastError());
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.stats.renderMs = ms;
    c.stats.pathsTraced = (long long)c.accW * c.accH * spp * nframes;
    return BTE_OK;
}

int renderFrames(int w, int h, int nframes) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (nframes <= 0) return fail(BTE_ERR_INVALID, "nframes must be positive");
    CHECK_RC(validateForRender(c));
    if (!c.dAccum.p || c.accW != w || c.accH != h) CHECK_RC(resetAccum(w, h));
    CHECK_RC(ensureAccel(c));
    CHECK_RC(c.mode == BTE_MODE_MEGAKERNEL ? renderMegakernel(nframes) : renderWavefront(nframes));
    // the reference denoises after the frame when its denoiser is on (src/render_kernel.cu:139-145)
    if (denoiser_active()) CHECK_RC(denoiser_present());
    return BTE_OK;
}

// ---------------------------------------------------------------------------
// multi-scene batches
// ---------------------------------------------------------------------------
void batchClear(Context& c) {
    c.batch.clear();
    c.batchPixels = 0;
    c.batchFrame = 0;
    ++c.batchGen;
    DBuf* bufs[] = {&c.dBatchScenes, &c.dBatchSlots, &c.dBatchChunk, &c.dBatchAccum, &c.dBatchFb,
                    &c.dBatchSum64, &c.dBatchAovAlb, &c.dBatchAovNrm, &c.dBatchDenoised,
                    &c.dBatchSegCount, &c.dBatchSegRows};
    for (DBuf* b : bufs) b->release();
    c.batchLayoutKey = -1;
    c.batchLastFrames = 0;
    c.batchDenoised = false;
}

// (Re)allocates and zeroes the batch's per-pixel outputs; frame index 0.
int batchResetAccum(Context& c) {
    const size_t N = (size_t)c.batchPixels;
    if (N == 0) return BTE_OK;
    CHECK_RC(c.dBatchAccum.ensure(N * sizeof(float4)));
    CHECK_RC(c.dBatchFb.ensure(N * sizeof(uint32_t)));
    HIPCHK(hipMemsetAsync(c.dBatchAccum.p, 0, N * sizeof(float4), c.stream));
    HIPCHK(hipMemsetAsync(c.dBatchFb.p, 0, N * sizeof(uint32_t), c.stream));
    if (c.aov) {
        CHECK_RC(c.dBatchAovAlb.ensure(N * sizeof(float4)));
        CHECK_RC(c.dBatchAovNrm.ensure(N * sizeof(float4)));
        HIPCHK(hipMemsetAsync(c.dBatchAovAlb.p, 0, N * sizeof(float4), c.stream));
        HIPCHK(hipMemsetAsync(c.dBatchAovNrm.p, 0, N * sizeof(float4), c.stream));
    }
    if (c.accum64) {
        CHECK_RC(c.dBatchSum64.ensure(N * 3 * sizeof(double)));
        HIPCHK(hipMemsetAsync(c.dBatchSum64.p, 0, N * 3 * sizeof(double), c.stream));
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    c.batchFrame = 0;
    c.batchDenoised = false;
    return BTE_OK;
}

// Light classes of a batch: the union over its scenes (a scene without a
// class simply makes no records of it).
int batchKinds(const Context& c, int& kindArea, int& kindSpot, int& kindEnv, int& catMask) {
    bool area = false, env = false, spots = false;
    catMask = 0;
    for (const BatchScene& b : c.batch) {
        area = area || b.data->emitterCount > 0;
        env = env || b.data->hasEnv;
        spots = spots || b.data->spotCount > 0;
        for (const BteBsdf& x : b.data->bsdfsHost) catMask |= 1 << x.type;
    }
    int k = 0;
    kindArea = area ? k++ : -1;
    kindSpot = spots ? k++ : -1;
    kindEnv = env ? k++ : -1;
    return k;
}

// Path-id layout of a pass of F frames: scene after scene, each range rounded
// up to kBatchChunk so that one byte per chunk names the scene.
long long batchPaddedPaths(const Context& c, int F) {
    long long pb = 0;
    for (const BatchScene& b : c.batch) pb += ((long long)b.w * b.h * F + kBatchChunk - 1) / kBatchChunk * kBatchChunk;
    return pb;
}

// First path id of batch slot `slot` in a pass of F frames (the layout above).
long long batchPathBase(const Context& c, int F, int slot) {
    long long pb = 0;
    for (int i = 0; i < slot; ++i)
        pb += ((long long)c.batch[i].w * c.batch[i].h * F + kBatchChunk - 1) / kBatchChunk * kBatchChunk;
    return pb;
}

int batchLayout(Context& c, int F, long long& paddedPaths, long long& slots, int& maxPaths, int& maxPixels) {
    const int S = (int)c.batch.size();
    std::vector<BatchSlot> tab((size_t)S);
    long long pb = 0, sb = 0;
    maxPaths = maxPixels = 0;
    for (int i = 0; i < S; ++i) {
        const long long px = (long long)c.batch[i].w * c.batch[i].h;
        tab[i] = BatchSlot{pb, sb, (int)(px * F), (int)px, (int)c.batch[i].pixelBase, 0};
        pb += (px * F + kBatchChunk - 1) / kBatchChunk * kBatchChunk;
        sb += px * F;
        maxPaths = std::max(maxPaths, (int)(px * F));
        maxPixels = std::max(maxPixels, (int)px);
    }
    paddedPaths = pb;
    slots = sb;
    if (c.batchLayoutKey != F || c.batchLayoutGen != c.batchGen) {
        std::vector<uint8_t> chunk((size_t)(pb / kBatchChunk), 0);
        for (int i = 0; i < S; ++i)
            for (long long k = tab[i].pathBase / kBatchChunk; k * kBatchChunk < tab[i].pathBase + tab[i].paths; ++k)
                chunk[(size_t)k] = (uint8_t)i;
        CHECK_RC(c.dBatchSlots.upload(tab.data(), tab.size() * sizeof(BatchSlot)));
        CHECK_RC(c.dBatchChunk.upload(chunk.data(), chunk.size()));
        c.batchLayoutKey = F;
        c.batchLayoutGen = c.batchGen;
    }
    return BTE_OK;
}

int renderBatch(int nframes) {
    Context& c = ctx();
    if (c.batch.empty()) return fail(BTE_ERR_STATE, "the batch has no scenes (bte_batch_add_scene)");
    if (nframes <= 0) return fail(BTE_ERR_INVALID, "nframes must be positive");
    if (c.mode == BTE_MODE_MEGAKERNEL) return fail(BTE_ERR_UNSUPPORTED, "batches run the wavefront path only");
    if (!c.dBatchAccum.p || (c.aov && !c.dBatchAovAlb.p) || (c.accum64 && !c.dBatchSum64.p))
        CHECK_RC(batchResetAccum(c));
    const int S = (int)c.batch.size();
    if (!c.dBatchScenes.p || c.dBatchScenes.bytes < (size_t)S * sizeof(DevScene) || c.batchLayoutGen != c.batchGen) {
        std::vector<DevScene> scenes((size_t)S);
        for (int i = 0; i < S; ++i) scenes[i] = makeScene(*c.batch[i].data, c.batch[i].w, c.batch[i].h);
        CHECK_RC(c.dBatchScenes.upload(scenes.data(), scenes.size() * sizeof(DevScene)));
    }
    int kindArea, kindSpot, kindEnv, catMask;
    const int kinds = batchKinds(c, kindArea, kindSpot, kindEnv, catMask);
    const int maxDepth = std::max(1, c.cfg.maxDepth);
    int F = c.cfg.framesInFlight;
    if (F <= 0) {
        const double budget = (double)(c.cfg.memoryBudgetMB > 0 ? c.cfg.memoryBudgetMB : 32768) * 1048576.0;
        F = (int)std::max(1.0, std::floor(budget / ((double)c.batchPixels * (double)bytesPerPath(kinds, kindSpot >= 0))));
    }
    F = std::max(1, std::min(F, nframes));
    // path ids (padded layout) occupy 30 bits of the path word
    while (F > 1 && batchPaddedPaths(c, F) > (1ll << 30)) --F;
    if (batchPaddedPaths(c, F) > (1ll << 30)) return fail(BTE_ERR_INVALID, "the batch exceeds 2^30 paths per frame");

    LaunchShape ls = query_launch_shape(c.device, c.cfg.traversalStack > 0 ? c.cfg.traversalStack : 16,
                                        c.cfg.traversalBlocksPerCU, c.cfg.workChunk, c.cfg.bvhWidth);
    for (const BatchScene& b : c.batch)
        if (ls.bvhWidth == 8 && !b.data->dNodes8.p) ls.bvhWidth = 4;   // a scene without a BVH8
    for (const BatchScene& b : c.batch)
        if (ls.bvhWidth == 4 && !b.data->dNodes4.p) ls.bvhWidth = 2;
    const bool useGraph = c.cfg.graphs != 0;
    const int tailAt = tailBounce(maxDepth);
    const size_t ctrInts = (size_t)kCtrStride * (size_t)(maxDepth + 1);
    std::vector<int> ctrHost(ctrInts);
    const size_t segInts = (size_t)(maxDepth + 1) * kSegQueues * (size_t)S;
    const size_t rowInts = (size_t)(maxDepth + 1) * kSegQueues * (size_t)(S + 1);
    CHECK_RC(c.dBatchSegCount.ensure(segInts * sizeof(int)));
    CHECK_RC(c.dBatchSegRows.ensure(rowInts * sizeof(int)));
    const DevScene* scenesDev = c.dBatchScenes.as<DevScene>();
    DevScene none;
    std::memset(&none, 0, sizeof none);
    double* sum64 = c.accum64 ? c.dBatchSum64.as<double>() : nullptr;

    c.stats = BteStats{};
    c.stats.framesInFlight = F;
    HIPCHK(hipEventRecord(c.ev0, c.stream));
    const long long graphLaunches0 = c.graphLaunches;
    int done = 0;
    while (done < nframes) {
        const int frames = std::min(F, nframes - done);
        long long padded = 0, slots = 0;
        int maxPaths = 0, maxPixels = 0;
        CHECK_RC(batchLayout(c, frames, padded, slots, maxPaths, maxPixels));
        WaveBuffers wb;
        std::memset(&wb, 0, sizeof wb);   // padding too: compared bytewise as graph keys
        CHECK_RC(ensureWaveBuffers(padded, kinds, kindSpot >= 0, maxDepth, wb));   // hit capture by padded path id
        if ((size_t)overflow_stack_ints(ls) * sizeof(int) > c.dOvf.bytes)
            return fail(BTE_ERR_STATE, "traversal overflow stack smaller than the launch shape needs");
        wb.kinds = kinds;
        wb.travCtr = nullptr;
        wb.catMask = catMask;
        wb.kindArea = kindArea;
        wb.kindSpot = kindSpot;
        wb.kindEnv = kindEnv;
        wb.grabRows = ls.grabRows;
        wb.aovAccAlb = c.aov ? c.dBatchAovAlb.as<float4>() : nullptr;
        wb.aovAccNrm = c.aov ? c.dBatchAovNrm.as<float4>() : nullptr;
        wb.sceneOfChunk = c.dBatchChunk.as<uint8_t>();
        wb.slots = c.dBatchSlots.as<BatchSlot>();
        wb.segCount = c.dBatchSegCount.as<int>();
        wb.segRows = c.dBatchSegRows.as<int>();
        wb.sceneCount = S;
        wb.paddedSlots = (int)padded;
        launch_fill_i32(c.dFrameBase.as<int>(), 1, (int)c.batchFrame, c.stream);
        constexpr int kSegment = 16;
        int b = 0;
        bool presented = false;
        for (int b0 = 0; b0 < maxDepth; b0 += kSegment) {
            const int b1 = std::min(b0 + kSegment, maxDepth);
            const bool tail = tailAt >= b0 && tailAt < b1;
            const int bEnd = tail ? tailAt : b1;
            const bool last = b1 == maxDepth || tail;
            std::vector<unsigned char> key;
            if (useGraph) {
                const char tag[8] = {'b', 'a', 't', 'c', 'h', 0, 0, 0};
                appendKey(key, tag);
                appendKey(key, c.batchGen);
                appendKey(key, wb);
                appendKey(key, ls);
                const long long params[] = {padded, slots, maxPaths, maxPixels, frames, b0, b1, maxDepth, tailAt,
                                            (long long)(uintptr_t)scenesDev, (long long)(uintptr_t)sum64,
                                            (long long)(uintptr_t)c.dBatchAccum.p, (long long)(uintptr_t)c.dBatchFb.p};
                appendKey(key, params);
            }
            CHECK_RC(runSegment(useGraph, key, [&]() {
                if (b0 == 0) {
                    launch_fill_i32(wb.counters, (long long)ctrInts, 0, c.stream);
                    launch_fill_i32(wb.work, (long long)kWorkStride * (maxDepth + 1), 0, c.stream);
                    launch_fill_i32(wb.segCount, (long long)segInts, 0, c.stream);
                    if (wb.hitCapture) launch_fill_i32(wb.hitCapture, padded * wb.maxRec, -2, c.stream);
                    launch_generate_batch(scenesDev, wb, maxPaths, slots, c.stream);
                }
                for (int bb = b0; bb < bEnd; ++bb) launch_bounce(none, wb, bb, maxDepth, ls, c.stream, nullptr, scenesDev);
                if (tail) launch_tail(none, wb, tailAt, maxDepth, ls, c.stream, nullptr, scenesDev);
                if (last)
                    launch_present_batch(wb, maxPixels, frames, c.dBatchAccum.as<float4>(), sum64,
                                         c.dBatchFb.as<uint32_t>(), c.stream);
            }));
            b = tail ? maxDepth : b1;
            if (last) {
                presented = true;
                break;
            }
            int next = 0;
            HIPCHK(hipMemcpyAsync(&next, wb.counters + (size_t)b1 * kCtrStride + kCtrRays, sizeof(int),
                                  hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (next == 0) break;
        }
        if (!presented)
            launch_present_batch(wb, maxPixels, frames, c.dBatchAccum.as<float4>(), sum64, c.dBatchFb.as<uint32_t>(),
                                 c.stream);
        HIPCHK(hipMemcpyAsync(ctrHost.data(), wb.counters, ctrInts * sizeof(int), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        HIPCHK(hipGetLastError());
        for (int i = 0; i < b; ++i) {
            const int* cb = ctrHost.data() + (size_t)i * kCtrStride;
            c.stats.extensionRays += cb[kCtrRays];
            c.stats.hits += cb[kCtrHits];
            for (int k = 0; k < kinds; ++k) c.stats.shadowRays += (k == kindSpot) ? cb[kCtrSpotRays] : cb[kCtrShadow + k];
            if (i < 64) c.stats.raysPerBounce[i] += cb[kCtrRays];
        }
        c.stats.bounces = std::max(c.stats.bounces, b);
        c.stats.pathsTraced += slots;
        c.batchFrame += (uint32_t)frames;
        c.batchLastFrames = frames;
        done += frames;
    }
    HIPCHK(hipEventRecord(c.ev1, c.stream));
    HIPCHK(hipEventSynchronize(c.ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.stats.renderMs = ms;
    c.stats.graphLaunches = (int)(c.graphLaunches - graphLaunches0);
    c.batchDenoised = false;
    return BTE_OK;
}

// Coordinate descent over the launch shape (BVH width, LDS stack, work chunk,
// traversal blocks per CU, tail-pass bounce).  timed(ms) renders one candidate's workload from a
// reset accumulation; each candidate runs twice (the first run captures its
// graphs) and keeps the second time.  Leaves the best configuration set.
template <class F>
int autotune(F&& timed, double* msBest) {
    Context& c = ctx();
    const BteConfig saved = c.cfg;
    const int savedInstr = c.instr;
    c.instr = 0;
    auto timeOne = [&](const BteConfig& cand, double& ms) -> int {
        c.cfg = cand;
        CHECK_RC(timed(ms));
        return timed(ms);
    };
    BteConfig best = saved;
    if (best.traversalStack == 0) best.traversalStack = 16;
    if (best.workChunk == 0) best.workChunk = 4;
    if (best.bvhWidth == 0) best.bvhWidth = 2;
    double bestMs = 0.0;
    int rc = timeOne(best, bestMs);
    auto tryCand = [&](BteConfig cand) {
        double ms = 0.0;
        if (rc == BTE_OK && (rc = timeOne(cand, ms)) == BTE_OK && ms < bestMs) {
            bestMs = ms;
            best = cand;
        }
    };
    const BteConfig startWidth = best;
    for (int w : {2, 4, 8}) {
        BteConfig cand = startWidth;
        cand.bvhWidth = w;
        if (w != startWidth.bvhWidth) tryCand(cand);
    }
    for (int st : {8, 16, 32}) {
        BteConfig cand = best;
        cand.traversalStack = st;
        if (st != best.traversalStack) tryCand(cand);
    }
    const BteConfig afterStack = best;
    for (int ch : {1, 2, 4, 8}) {
        BteConfig cand = afterStack;
        cand.workChunk = ch;
        if (ch != afterStack.workChunk) tryCand(cand);
    }
    int cus = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus <= 0) cus = 1;
    const int occ = std::max(
        1, query_launch_shape(c.device, best.traversalStack, 0, best.workChunk, best.bvhWidth).extendBlocks / cus);
    const BteConfig afterChunk = best;
    for (int bpc : {0, (3 * occ) / 4, occ / 2}) {
        BteConfig cand = afterChunk;
        cand.traversalBlocksPerCU = bpc;
        if (bpc != afterChunk.traversalBlocksPerCU && (bpc == 0 || bpc < occ)) tryCand(cand);
    }
    const BteConfig afterBlocks = best;
    for (int tb : {0, 2, 3, 4, 5, 6, 8}) {
        BteConfig cand = afterBlocks;
        cand.tailBounce = tb;
        if (tb != afterBlocks.tailBounce && tb < std::max(1, afterBlocks.maxDepth)) tryCand(cand);
    }
    c.instr = savedInstr;
    if (rc != BTE_OK) {
        c.cfg = saved;
        return rc;
    }
    c.cfg = best;
    if (msBest) *msBest = bestMs;
    return BTE_OK;
}

int batchSlot(const Context& c, int slot) {
    if (slot < 0 || slot >= (int)c.batch.size())
        return fail(BTE_ERR_INVALID, "batch slot %d out of range (%d scenes)", slot, (int)c.batch.size());
    return BTE_OK;
}

}  // namespace

int ensure_device() { return ensureDevice(); }
hipStream_t main_stream() { return ctx().stream; }

int run_captured(const std::vector<unsigned char>& key, const std::function<void()>& emit) {
    Context& c = ctx();
    return runSegment(c.cfg.graphs != 0 && c.instr == 0, key, emit);
}


int render_view(RenderView& v) {
    CHECK_RC(ensureDevice());
    Context& c = ctx();
    if (!c.dAccum.p || c.accW <= 0) return fail(BTE_ERR_STATE, "no accumulation buffer (render first)");
    v.stream = c.stream;
    v.device = c.device;
    v.width = c.accW;
    v.height = c.accH;
    v.accum = c.dAccum.as<float4>();
    v.aovAlb = (c.aov && c.dAovAccAlb.p) ? c.dAovAccAlb.as<float4>() : nullptr;
    v.aovNrm = (c.aov && c.dAovAccNrm.p) ? c.dAovAccNrm.as<float4>() : nullptr;
    v.fb = c.externalFb ? c.externalFb : c.dFb.as<uint32_t>();
    v.frameIndex = c.frameIndex;
    return BTE_OK;
}

}  // namespace bte

using namespace bte;

extern "C" {

const char* bte_last_error(void) { return g_err; }

int bte_init_scene(const BteTriangle* triangles, int triangleCount, const BteFloat3* materials, int materialCount,
                   const int* triangleMaterialIds, int imageWidth, int imageHeight) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (triangleCount < 0 || (triangleCount > 0 && !triangles))
        return fail(BTE_ERR_INVALID, "bte_init_scene: invalid triangle array");
    if (triangleCount >= (1 << 28)) return fail(BTE_ERR_INVALID, "bte_init_scene: at most 2^28 triangles");
    c.tris.assign(triangles, triangles + triangleCount);
    c.materials.assign(materials ? materials : nullptr, materials ? materials + std::max(0, materialCount) : nullptr);
    if (triangleMaterialIds) c.triMaterialIds.assign(triangleMaterialIds, triangleMaterialIds + triangleCount);
    else c.triMaterialIds.assign((size_t)triangleCount, 0);
    CHECK_RC(c.dTris.upload(c.tris.data(), c.tris.size() * sizeof(BteTriangle)));
    CHECK_RC(c.dMaterials.upload(c.materials.data(), c.materials.size() * sizeof(BteFloat3)));
    CHECK_RC(c.dTriMat.upload(c.triMaterialIds.data(), c.triMaterialIds.size() * sizeof(int)));
    c.sceneW = imageWidth;
    c.sceneH = imageHeight;
    c.bvhDirty = true;
    return rebuildEmitterOfTri();
}

int bte_init(const BteTriangle* tri, BteFloat3 color, int imageWidth, int imageHeight) {
    if (!tri) return fail(BTE_ERR_INVALID, "bte_init: null triangle");
    const int matId = 0;
    return bte_init_scene(tri, 1, &color, 1, &matId, imageWidth, imageHeight);
}

int bte_init_camera(const BteCamera* camera) {
    if (!camera) return fail(BTE_ERR_INVALID, "bte_init_camera: null camera");
    ctx().camera = *camera;
    ctx().haveCamera = true;
    return BTE_OK;
}

int bte_init_triangle_emission(const BteFloat3* triangleEmission, int triangleCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (!triangleEmission || triangleCount <= 0) {
        c.haveEmission = false;
        c.dEmission.release();
        return BTE_OK;
    }
    CHECK_RC(c.dEmission.upload(triangleEmission, (size_t)triangleCount * sizeof(BteFloat3)));
    c.haveEmission = true;
    return BTE_OK;
}

int bte_init_triangle_emitter_flags(const uint8_t* flags, int triangleCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (!flags || triangleCount <= 0) {
        c.haveFlags = false;
        c.dFlags.release();
        return BTE_OK;
    }
    CHECK_RC(c.dFlags.upload(flags, (size_t)triangleCount));
    c.haveFlags = true;
    return BTE_OK;
}

int bte_init_emitter_table(const BteEmitter* emitters, int emitterCount, const int* emitterTriIndices,
                           int emitterTriIndexCount, const float* emitterTriCdf, int emitterTriCdfCount,
                           const float* sceneEmitterCdf) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (!emitters || emitterCount <= 0) {
        c.emitterCount = 0;
        c.emittersHost.clear();
        c.emTriIdxHost.clear();
        c.emTriIdxCount = c.emTriCdfCount = 0;
        c.dEmitters.release(); c.dEmTriIdx.release(); c.dEmTriCdf.release(); c.dSceneCdf.release();
        c.dEmitterOfTri.release();
        return BTE_OK;
    }
    if (!emitterTriIndices || !emitterTriCdf || !sceneEmitterCdf || emitterTriIndexCount <= 0 || emitterTriCdfCount <= 0)
        return fail(BTE_ERR_INVALID, "bte_init_emitter_table: missing emitter tables");
    c.emittersHost.assign(emitters, emitters + emitterCount);
    c.emTriIdxHost.assign(emitterTriIndices, emitterTriIndices + emitterTriIndexCount);
    c.emitterCount = emitterCount;
    c.emTriIdxCount = emitterTriIndexCount;
    c.emTriCdfCount = emitterTriCdfCount;
    CHECK_RC(c.dEmitters.upload(emitters, (size_t)emitterCount * sizeof(BteEmitter)));
    CHECK_RC(c.dEmTriIdx.upload(emitterTriIndices, (size_t)emitterTriIndexCount * sizeof(int)));
    CHECK_RC(c.dEmTriCdf.upload(emitterTriCdf, (size_t)emitterTriCdfCount * sizeof(float)));
    CHECK_RC(c.dSceneCdf.upload(sceneEmitterCdf, (size_t)(emitterCount + 1) * sizeof(float)));
    return rebuildEmitterOfTri();
}

int bte_init_bsdfs(const BteBsdf* bsdfs, int bsdfCount, const int* triangleBsdfIds, int triangleCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (!bsdfs || bsdfCount <= 0 || !triangleBsdfIds || triangleCount <= 0) {
        c.bsdfsHost.clear();
        c.triBsdfHost.clear();
        c.dBsdfs.release();
        c.dTriBsdf.release();
        return BTE_OK;
    }
    for (int i = 0; i < bsdfCount; ++i)
        if (bsdfs[i].type < BTE_BSDF_DIFFUSE || bsdfs[i].type > BTE_BSDF_DISNEY)
            return fail(BTE_ERR_INVALID, "bte_init_bsdfs: unsupported bsdf type %d at %d", bsdfs[i].type, i);
    for (int i = 0; i < triangleCount; ++i)
        if (triangleBsdfIds[i] < 0 || triangleBsdfIds[i] >= bsdfCount)
            return fail(BTE_ERR_INVALID, "bte_init_bsdfs: triangle %d references bsdf %d of %d", i, triangleBsdfIds[i],
                        bsdfCount);
    c.bsdfsHost.assign(bsdfs, bsdfs + bsdfCount);
    c.triBsdfHost.assign(triangleBsdfIds, triangleBsdfIds + triangleCount);
    std::vector<float> packed((size_t)bsdfCount * 16);
    for (int i = 0; i < bsdfCount; ++i) {
        const BteBsdf& b = bsdfs[i];
        float* p = &packed[(size_t)i * 16];
        std::memcpy(p + 0, &b.p0, 16);
        std::memcpy(p + 4, &b.p1, 16);
        std::memcpy(p + 8, &b.p2, 16);
        std::memcpy(p + 12, &b.p3, 12);
        std::memcpy(p + 15, &b.type, 4);
    }
    CHECK_RC(c.dBsdfs.upload(packed.data(), packed.size() * sizeof(float)));
    CHECK_RC(c.dTriBsdf.upload(triangleBsdfIds, (size_t)triangleCount * sizeof(int)));
    return BTE_OK;
}

int bte_init_spotlights(const BteSpotlight* spotlights, int spotlightCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (!spotlights || spotlightCount <= 0) {
        c.spotCount = 0;
        c.dSpots.release();
        return BTE_OK;
    }
    CHECK_RC(c.dSpots.upload(spotlights, (size_t)spotlightCount * sizeof(BteSpotlight)));
    c.spotCount = spotlightCount;
    return BTE_OK;
}

// One texture set (src/host/scene_upload.cu:213-271): entry i serves material
// i; a null entry has no texture.  The RGBA8 texels of all entries go back to
// back into one buffer, with (offset, width, height) per entry.
static int uploadTextureSet(const uint8_t* const* pixels, const int* widths, const int* heights, int count,
                            DBuf& texels, DBuf& info, int& outCount, const char* what) {
    texels.release();
    info.release();
    outCount = 0;
    if (!pixels || count <= 0) return BTE_OK;
    std::vector<int> tab((size_t)count * 4, 0);
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        if (!pixels[i]) continue;
        if (!widths || !heights || widths[i] <= 0 || heights[i] <= 0 || widths[i] > 65536 || heights[i] > 65536)
            return fail(BTE_ERR_INVALID, "%s %d: invalid size", what, i);
        tab[4 * (size_t)i] = (int)total;
        tab[4 * (size_t)i + 1] = widths[i];
        tab[4 * (size_t)i + 2] = heights[i];
        total += (size_t)widths[i] * (size_t)heights[i];
        if (total >= (1ull << 31)) return fail(BTE_ERR_INVALID, "%s: more than 2^31 texels", what);
    }
    std::vector<uint32_t> all(std::max<size_t>(total, 1), 0u);
    for (int i = 0; i < count; ++i)
        if (pixels[i])
            std::memcpy(all.data() + tab[4 * (size_t)i], pixels[i], (size_t)widths[i] * (size_t)heights[i] * 4);
    CHECK_RC(texels.upload(all.data(), all.size() * sizeof(uint32_t)));
    CHECK_RC(info.upload(tab.data(), tab.size() * sizeof(int)));
    outCount = count;
    return BTE_OK;
}
int bte_init_textures(const uint8_t* const* pixels, const int* widths, const int* heights, int textureCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    return uploadTextureSet(pixels, widths, heights, textureCount, c.dTexels, c.dTexInfo, c.texCount, "texture");
}
int bte_init_mr_textures(const uint8_t* const* pixels, const int* widths, const int* heights, int textureCount) {
    ENSURE_DEVICE();
    Context& c = ctx();
    return uploadTextureSet(pixels, widths, heights, textureCount, c.dMrTexels, c.dMrInfo, c.mrCount,
                            "metallic-roughness map");
}

int bte_init_env_map(const float* hdriPixelsRGBA, int width, int height, float intensity, float rotationDeg) {
    ENSURE_DEVICE();
    Context& c = ctx();
    c.hasEnv = false;
    c.dEnv.release(); c.dEnvMarg.release(); c.dEnvCond.release(); c.dEnvRowSums.release(); c.dEnvTotal.release();
    c.dEnvMargGuide.release(); c.dEnvCondGuide.release(); c.dEnvMono.release(); c.dEnvRec.release();
    c.envGuides = false;
    c.envW = c.envH = 0;
    c.envTotal = 0.0f;
    if (!hdriPixelsRGBA || width <= 0 || height <= 0) return BTE_OK;
    const size_t texels = (size_t)width * (size_t)height;
    CHECK_RC(c.dEnv.upload(hdriPixelsRGBA, texels * sizeof(float4)));
    CHECK_RC(c.dEnvMarg.ensure((size_t)(height + 1) * sizeof(float)));
    CHECK_RC(c.dEnvCond.ensure((size_t)height * (size_t)(width + 1) * sizeof(float)));
    CHECK_RC(c.dEnvRowSums.ensure((size_t)height * sizeof(double)));
    CHECK_RC(c.dEnvTotal.ensure(sizeof(float)));
    CHECK_RC(c.dEnvMargGuide.ensure((size_t)(height + 1) * sizeof(int)));
    CHECK_RC(c.dEnvCondGuide.ensure((size_t)height * (size_t)(width + 1) * sizeof(int)));
    CHECK_RC(c.dEnvMono.ensure(sizeof(int)));
    launch_env_cdf(c.dEnv.as<float4>(), width, height, c.dEnvMarg.as<float>(), c.dEnvCond.as<float>(),
                   c.dEnvRowSums.as<double>(), c.dEnvTotal.as<float>(), c.dEnvMargGuide.as<int>(),
                   c.dEnvCondGuide.as<int>(), c.dEnvMono.as<int>(), c.stream);
    float total = 0.0f;
    int monotone = 0;
    HIPCHK(hipMemcpyAsync(&total, c.dEnvTotal.p, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(&monotone, c.dEnvMono.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    c.envGuides = monotone != 0;
    CHECK_RC(c.dEnvRec.ensure(env_record_bytes(width, height)));
    launch_env_records(c.dEnv.as<float4>(), c.dEnvMarg.as<float>(), c.dEnvCond.as<float>(), width, height, total,
                       c.dEnvRec.as<float4>(), c.stream);
    HIPCHK(hipGetLastError());
    c.envW = width;
    c.envH = height;
    c.envIntensity = intensity;
    c.envRot = rotationDeg * (3.14159265358979323846f / 180.0f);
    c.envTotal = total;
    c.hasEnv = true;
    return BTE_OK;
}

int bte_build_accel(void) {
    ENSURE_DEVICE();
    ctx().bvhDirty = true;
    return ensureAccel(ctx());
}

int bte_register_framebuffer(uint32_t* deviceRGBA8) {
    ctx().externalFb = deviceRGBA8;
    return BTE_OK;
}

int bte_reset_accumulation(int imageWidth, int imageHeight) {
    ENSURE_DEVICE();
    return resetAccum(imageWidth, imageHeight);
}

int bte_render(int imageWidth, int imageHeight) { return renderFrames(imageWidth, imageHeight, 1); }

int bte_render_frames(int imageWidth, int imageHeight, int nframes) {
    return renderFrames(imageWidth, imageHeight, nframes);
}

int bte_set_render_mode(int mode) {
    if (mode != BTE_MODE_MEGAKERNEL && mode != BTE_MODE_WAVEFRONT) return fail(BTE_ERR_INVALID, "unknown mode %d", mode);
    Context& c = ctx();
    if (mode == c.mode) return BTE_OK;
    c.mode = mode;
    if (c.accW > 0) return bte_reset_accumulation(c.accW, c.accH);   // render_kernel.cu:171-176
    return BTE_OK;
}
int bte_get_render_mode(void) { return ctx().mode; }

int bte_cleanup(void) {
    Context& c = ctx();
    if (c.deviceReady) (void)hipStreamSynchronize(c.stream);
    dropGraphs();
    denoiser_release();
    batchClear(c);
    c.dFrameBase.release();
    c.dWork.release();
    c.dAovAlb.release(); c.dAovNrm.release(); c.dAovAccAlb.release(); c.dAovAccNrm.release();
    c.aov = false;
    c.dTexels.release(); c.dTexInfo.release(); c.dMrTexels.release(); c.dMrInfo.release();
    c.texCount = c.mrCount = 0;
    DBuf* bufs[] = {&c.dTris, &c.dMaterials, &c.dTriMat, &c.dTriBsdf, &c.dBsdfs, &c.dEmission, &c.dFlags,
                    &c.dEmitterOfTri, &c.dEmitters, &c.dEmTriIdx, &c.dEmTriCdf, &c.dSceneCdf, &c.dSpots, &c.dNodes, &c.dNodes4, &c.dNodes8,
                    &c.dLeafTris, &c.dEnv, &c.dEnvMarg, &c.dEnvCond, &c.dEnvRowSums, &c.dEnvTotal, &c.dEnvMargGuide,
                    &c.dEnvCondGuide, &c.dEnvMono, &c.dEnvRec, &c.dAccum,
                    &c.dFb, &c.dSum64, &c.dHit, &c.dCatQ, &c.dCat, &c.dFinal, &c.dShO, &c.dShD, &c.dShC, &c.dShX, &c.dCounters, &c.dHitCap,
                    &c.dOvf};
    for (DBuf* b : bufs) b->release();
    c.dTrav.release();
    for (int i = 0; i < 2; ++i) {
        c.dRayO[i].release(); c.dRayD[i].release(); c.dAux[i].release();
    }
    c.tris.clear(); c.materials.clear(); c.triMaterialIds.clear(); c.bsdfsHost.clear(); c.triBsdfHost.clear();
    c.emittersHost.clear(); c.emTriIdxHost.clear();
    c.emitterCount = c.spotCount = 0;
    c.haveEmission = c.haveFlags = c.haveCamera = c.hasEnv = false;
    c.bvhDirty = true;
    c.accW = c.accH = 0;
    c.frameIndex = 0;
    c.cap = 0;
    c.externalFb = nullptr;
    return BTE_OK;
}

int bte_set_accel_builder(int builder) {
    if (builder != 0 && builder != 1) return fail(BTE_ERR_INVALID, "accel builder must be 0 (host) or 1 (GPU)");
    Context& c = ctx();
    if (c.accelBuilder != builder) c.bvhDirty = true;   // the next render builds with the new builder
    c.accelBuilder = builder;
    return BTE_OK;
}

int bte_get_accel_stats(BteAccelStats* stats) {
    if (!stats) return fail(BTE_ERR_INVALID, "null pointer");
    *stats = ctx().lastAccel;
    return BTE_OK;
}

int bte_set_config(const BteConfig* config) {
    if (!config) return fail(BTE_ERR_INVALID, "null config");
    if (config->maxDepth <= 0 || config->maxDepth > 4096) return fail(BTE_ERR_INVALID, "maxDepth out of range");
    if (config->megakernelSpp <= 0) return fail(BTE_ERR_INVALID, "megakernelSpp must be positive");
    if (config->framesInFlight < 0 || config->memoryBudgetMB < 0) return fail(BTE_ERR_INVALID, "negative config");
    if (config->traversalStack != 0 && config->traversalStack != 8 && config->traversalStack != 16 &&
        config->traversalStack != 32)
        return fail(BTE_ERR_INVALID, "traversalStack must be 0, 8, 16 or 32");
    if (config->graphs != 0 && config->graphs != 1) return fail(BTE_ERR_INVALID, "graphs must be 0 or 1");
    if (config->workChunk != 0 && config->workChunk != 1 && config->workChunk != 2 && config->workChunk != 4 &&
        config->workChunk != 8)
        return fail(BTE_ERR_INVALID, "workChunk must be 0, 1, 2, 4 or 8");
    if (config->traversalBlocksPerCU < 0 || config->traversalBlocksPerCU > 64)
        return fail(BTE_ERR_INVALID, "traversalBlocksPerCU must be in [0, 64]");
    if (config->bvhWidth != 0 && config->bvhWidth != 2 && config->bvhWidth != 4 && config->bvhWidth != 8)
        return fail(BTE_ERR_INVALID, "bvhWidth must be 0, 2, 4 or 8");
    if (config->tailBounce < 0) return fail(BTE_ERR_INVALID, "tailBounce must not be negative");
    ctx().cfg = *config;
    return BTE_OK;
}
int bte_get_default_config(BteConfig* config) {
    if (!config) return fail(BTE_ERR_INVALID, "null config");
    *config = kDefaultConfig;
    return BTE_OK;
}
int bte_get_config(BteConfig* config) {
    if (!config) return fail(BTE_ERR_INVALID, "null config");
    *config = ctx().cfg;
    return BTE_OK;
}
int bte_autotune(int width, int height, int frames, double* msBest) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (width <= 0 || height <= 0 || frames <= 0) return fail(BTE_ERR_INVALID, "autotune needs positive sizes");
    if (c.mode != BTE_MODE_WAVEFRONT) return fail(BTE_ERR_STATE, "autotune applies to the wavefront mode");
    CHECK_RC(autotune([&](double& ms) -> int {
        CHECK_RC(bte_reset_accumulation(width, height));
        CHECK_RC(renderFrames(width, height, frames));
        ms = c.stats.renderMs;
        return BTE_OK;
    }, msBest));
    return bte_reset_accumulation(width, height);
}

int bte_batch_autotune(int frames, double* msBest) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (frames <= 0) return fail(BTE_ERR_INVALID, "autotune needs a positive frame count");
    if (c.batch.empty()) return fail(BTE_ERR_STATE, "the batch has no scenes");
    if (c.mode != BTE_MODE_WAVEFRONT) return fail(BTE_ERR_STATE, "autotune applies to the wavefront mode");
    CHECK_RC(autotune([&](double& ms) -> int {
        CHECK_RC(batchResetAccum(c));
        CHECK_RC(renderBatch(frames));
        ms = c.stats.renderMs;
        return BTE_OK;
    }, msBest));
    return batchResetAccum(c);
}

int bte_get_stats(BteStats* stats) {
    if (!stats) return fail(BTE_ERR_INVALID, "null stats");
    *stats = ctx().stats;
    return BTE_OK;
}
int bte_frame_index(void) { return (int)ctx().frameIndex; }
int bte_set_frame_index(int frameIndex) {
    if (frameIndex < 0) return fail(BTE_ERR_INVALID, "frame index must be non-negative");
    ctx().frameIndex = (uint32_t)frameIndex;
    return BTE_OK;
}
int bte_set_instrumentation(int flags) {
    if (flags < 0 || flags > 3) return fail(BTE_ERR_INVALID, "unknown instrumentation flags %d", flags);
    ctx().instr = flags;
    return BTE_OK;
}

int bte_read_framebuffer(uint32_t* hostRGBA8, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t N = (size_t)c.accW * c.accH;
    if (!hostRGBA8 || count < N) return fail(BTE_ERR_INVALID, "framebuffer copy needs %zu pixels", N);
    const void* src = c.externalFb ? (const void*)c.externalFb : c.dFb.p;
    if (!src) return fail(BTE_ERR_STATE, "nothing rendered yet");
    HIPCHK(hipMemcpy(hostRGBA8, src, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_read_accum(float* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t N = (size_t)c.accW * c.accH;
    if (!hostRGB || count < 3 * N) return fail(BTE_ERR_INVALID, "accum copy needs %zu floats", 3 * N);
    if (!c.dAccum.p) return fail(BTE_ERR_STATE, "no accumulation buffer");
    std::vector<float> tmp(4 * N);
    HIPCHK(hipMemcpy(tmp.data(), c.dAccum.p, N * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

int bte_enable_aov(int enable) {
    ENSURE_DEVICE();
    Context& c = ctx();
    c.aov = enable != 0;
    if (c.aov && c.accW > 0) {
        const size_t N = (size_t)c.accW * c.accH;
        CHECK_RC(c.dAovAccAlb.ensure(N * sizeof(float4)));
        CHECK_RC(c.dAovAccNrm.ensure(N * sizeof(float4)));
        HIPCHK(hipMemset(c.dAovAccAlb.p, 0, N * sizeof(float4)));
        HIPCHK(hipMemset(c.dAovAccNrm.p, 0, N * sizeof(float4)));
    }
    return BTE_OK;
}

int bte_read_aov(float* hostAlbedo, float* hostNormal, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t N = (size_t)c.accW * c.accH;
    if (count < 3 * N) return fail(BTE_ERR_INVALID, "AOV copy needs %zu floats per buffer", 3 * N);
    if (!c.aov || !c.dAovAccAlb.p) return fail(BTE_ERR_STATE, "AOVs are not enabled");
    std::vector<float> tmp(4 * N);
    float* dst[2] = {hostAlbedo, hostNormal};
    const DBuf* src[2] = {&c.dAovAccAlb, &c.dAovAccNrm};
    for (int k = 0; k < 2; ++k) {
        if (!dst[k]) continue;
        HIPCHK(hipMemcpy(tmp.data(), src[k]->p, N * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; ++i) {
            dst[k][3 * i] = tmp[4 * i];
            dst[k][3 * i + 1] = tmp[4 * i + 1];
            dst[k][3 * i + 2] = tmp[4 * i + 2];
        }
    }
    return BTE_OK;
}

int bte_enable_accum64(int enable) {
    ENSURE_DEVICE();
    Context& c = ctx();
    c.accum64 = enable != 0;
    if (c.accum64 && c.accW > 0) {
        const size_t N = (size_t)c.accW * c.accH;
        CHECK_RC(c.dSum64.ensure(N * 3 * sizeof(double)));
        HIPCHK(hipMemset(c.dSum64.p, 0, N * 3 * sizeof(double)));
    }
    return BTE_OK;
}

int bte_read_accum64(double* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t N = (size_t)c.accW * c.accH;
    if (!hostRGB || count < 3 * N) return fail(BTE_ERR_INVALID, "accum64 copy needs %zu doubles", 3 * N);
    if (!c.accum64 || !c.dSum64.p) return fail(BTE_ERR_STATE, "binary64 accumulation not enabled");
    HIPCHK(hipMemcpy(hostRGB, c.dSum64.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_accum_device_ptr(void** devicePtr) {
    if (!devicePtr) return fail(BTE_ERR_INVALID, "null pointer");
    *devicePtr = ctx().dAccum.p;
    return BTE_OK;
}

int bte_read_sample_radiance(float* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t P = (size_t)c.lastPassPaths;
    if (!hostRGB || count < 3 * P) return fail(BTE_ERR_INVALID, "sample copy needs %zu floats", 3 * P);
    if (P == 0) return fail(BTE_ERR_STATE, "no wavefront pass recorded");
    std::vector<float> tmp(4 * P);
    HIPCHK(hipMemcpy(tmp.data(), c.dFinal.p, P * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

int bte_set_hit_capture(int maxRecord) {
    if (maxRecord < 0 || maxRecord > 64) return fail(BTE_ERR_INVALID, "maxRecord must be in [0, 64]");
    ctx().hitCaptureRec = maxRecord;
    return BTE_OK;
}

int bte_read_hit_capture(int* hostTris, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    const size_t n = (size_t)c.lastPassPaths * (size_t)c.hitCaptureRec;
    if (!hostTris || count < n) return fail(BTE_ERR_INVALID, "hit capture copy needs %zu ints", n);
    if (n == 0 || !c.dHitCap.p) return fail(BTE_ERR_STATE, "no hit capture recorded");
    HIPCHK(hipMemcpy(hostTris, c.dHitCap.p, n * sizeof(int), hipMemcpyDeviceToHost));
    return BTE_OK;
}

static int readEnvCdf(const SceneData& c, float* marginal, float* conditional, float* totalPower) {
    if (!c.hasEnv) return fail(BTE_ERR_STATE, "no environment map");
    if (marginal) HIPCHK(hipMemcpy(marginal, c.dEnvMarg.p, (size_t)(c.envH + 1) * sizeof(float), hipMemcpyDeviceToHost));
    if (conditional)
        HIPCHK(hipMemcpy(conditional, c.dEnvCond.p, (size_t)c.envH * (size_t)(c.envW + 1) * sizeof(float),
                         hipMemcpyDeviceToHost));
    if (totalPower) *totalPower = c.envTotal;
    return BTE_OK;
}

int bte_read_env_cdf(float* marginal, float* conditional, float* totalPower) {
    ENSURE_DEVICE();
    return readEnvCdf(ctx(), marginal, conditional, totalPower);
}

int bte_batch_read_env_cdf(int slot, float* marginal, float* conditional, float* totalPower) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    return readEnvCdf(*c.batch[slot].data, marginal, conditional, totalPower);
}

int bte_trace_rays(const float* rays, const float* tmaxPerRay, long long n, float tmin, float tmax, int anyHit,
                   int* tri, float* t, float* u, float* v) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (n <= 0) return BTE_OK;
    if (!rays || !tri || (!anyHit && (!t || !u || !v))) return fail(BTE_ERR_INVALID, "bte_trace_rays: null buffer");
    if (n >= (1ll << 31) - 128) return fail(BTE_ERR_INVALID, "bte_trace_rays: too many rays");
    if (c.tris.empty()) return fail(BTE_ERR_STATE, "no triangles uploaded");
    CHECK_RC(ensureAccel(c));
    if (c.accW <= 0) { c.accW = std::max(1, c.sceneW); c.accH = std::max(1, c.sceneH); }
    CHECK_RC(c.dOvf.ensure((size_t)overflow_stack_ints_max(c.device) * sizeof(int)));
    DBuf dr, dtm, dtri, dt, du, dv;
    CHECK_RC(dr.upload(rays, (size_t)n * 6 * sizeof(float)));
    if (anyHit && tmaxPerRay) CHECK_RC(dtm.upload(tmaxPerRay, (size_t)n * sizeof(float)));
    CHECK_RC(dtri.ensure((size_t)n * sizeof(int)));
    CHECK_RC(dt.ensure((size_t)n * sizeof(float)));
    CHECK_RC(du.ensure((size_t)n * sizeof(float)));
    CHECK_RC(dv.ensure((size_t)n * sizeof(float)));
    const DevScene sc = makeScene(c, c.accW, c.accH);
    launch_trace_rays(sc, dr.as<float>(), (anyHit && tmaxPerRay) ? dtm.as<float>() : nullptr, n, tmin, tmax, anyHit,
                      dtri.as<int>(), dt.as<float>(), du.as<float>(), dv.as<float>(), c.dOvf.as<int>(), c.stream);
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(tri, dtri.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
    if (!anyHit) {
        HIPCHK(hipMemcpy(t, dt.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(u, du.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(v, dv.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    }
    return BTE_OK;
}

int bte_trace_rays_noalloc(const float* rays, const float* tmaxPerRay, long long n, float tmin, float tmax,
                           int anyHit, int* tri, float* t, float* u, float* v) {
    Context& c = ctx();
    // no set-up here: the device, the scene and its BVH come from earlier calls
    if (!c.deviceReady) return fail(BTE_ERR_STATE, "bte_trace_rays_noalloc: no device set up (upload a scene first)");
    if (n <= 0) return BTE_OK;
    if (!rays || !tri || (!anyHit && (!t || !u || !v)))
        return fail(BTE_ERR_INVALID, "bte_trace_rays_noalloc: null buffer");
    constexpr long long kMaxRays = (1ll << 31) - (1ll << 20);   // the work counter is a 32-bit int
    if (n > kMaxRays)
        return fail(BTE_ERR_INVALID, "bte_trace_rays_noalloc: %lld rays exceed the maximum of %lld per call", n,
                    kMaxRays);
    if (c.tris.empty()) return fail(BTE_ERR_STATE, "bte_trace_rays_noalloc: no triangles uploaded");
    if (c.bvhDirty)
        return fail(BTE_ERR_STATE, "bte_trace_rays_noalloc: the BVH is not built (call bte_build_accel first; "
                                   "this entry point builds nothing)");
    if (c.bvhDepth > noalloc_max_bvh_depth())
        return fail(BTE_ERR_UNSUPPORTED, "bte_trace_rays_noalloc: BVH depth %d exceeds the %d levels its fixed "
                                         "traversal stacks hold", c.bvhDepth, noalloc_max_bvh_depth());
    const void* bufs[] = {rays, (anyHit && tmaxPerRay) ? tmaxPerRay : nullptr, tri, anyHit ? nullptr : t,
                          anyHit ? nullptr : u, anyHit ? nullptr : v};
    for (const void* b : bufs) {
        if (!b) continue;
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, b) != hipSuccess) {
            (void)hipGetLastError();
            return fail(BTE_ERR_INVALID, "bte_trace_rays_noalloc: buffer %p is not device memory", b);
        }
        if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != c.device)
            return fail(BTE_ERR_INVALID, "bte_trace_rays_noalloc: buffer %p is not memory of device %d", b, c.device);
    }
    int* scr = nullptr;
    const hipError_t e = noalloc_scratch(&scr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(BTE_ERR_DEVICE, "bte_trace_rays_noalloc: scratch address lookup failed: %s", hipGetErrorString(e));
    }
    const DevScene sc = makeScene(c, std::max(1, c.sceneW), std::max(1, c.sceneH));
    launch_trace_rays_noalloc(sc, rays, (anyHit && tmaxPerRay) ? tmaxPerRay : nullptr, n, tmin, tmax, anyHit, tri,
                              t, u, v, scr, c.stream);
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// ---- multi-scene batches -------------------------------------------------
int bte_batch_add_scene(int width, int height) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (width <= 0 || height <= 0) return fail(BTE_ERR_INVALID, "invalid resolution %dx%d", width, height);
    if ((long long)width * height > (1ll << 30)) return fail(BTE_ERR_INVALID, "more than 2^30 pixels");
    if (c.batch.size() >= 255) return fail(BTE_ERR_UNSUPPORTED, "a batch holds at most 255 scenes");
    CHECK_RC(validateForRender(c));
    CHECK_RC(ensureAccel(c));
    HIPCHK(hipStreamSynchronize(c.stream));
    dropGraphs();
    BatchScene b;
    b.data = std::make_unique<SceneData>(std::move(static_cast<SceneData&>(c)));
    static_cast<SceneData&>(c) = SceneData();
    b.w = width;
    b.h = height;
    b.pixelBase = c.batchPixels;
    c.batchPixels += (long long)width * height;
    c.batch.push_back(std::move(b));
    ++c.batchGen;
    CHECK_RC(batchResetAccum(c));
    return (int)c.batch.size() - 1;
}

int bte_batch_scene_count(void) { return (int)ctx().batch.size(); }

int bte_batch_render(int nframes) {
    ENSURE_DEVICE();
    CHECK_RC(renderBatch(nframes));
    return BTE_OK;
}

int bte_batch_reset(void) {
    ENSURE_DEVICE();
    return batchResetAccum(ctx());
}

int bte_batch_clear(void) {
    ENSURE_DEVICE();
    Context& c = ctx();
    HIPCHK(hipStreamSynchronize(c.stream));
    dropGraphs();
    batchClear(c);
    return BTE_OK;
}

int bte_batch_frame_index(void) { return (int)ctx().batchFrame; }

int bte_batch_read_accum(int slot, float* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    const BatchScene& b = c.batch[slot];
    const size_t N = (size_t)b.w * b.h;
    if (!hostRGB || count < 3 * N) return fail(BTE_ERR_INVALID, "accum copy needs %zu floats", 3 * N);
    if (!c.dBatchAccum.p) return fail(BTE_ERR_STATE, "no batch accumulation (render first)");
    std::vector<float> tmp(4 * N);
    HIPCHK(hipMemcpy(tmp.data(), c.dBatchAccum.as<float4>() + b.pixelBase, N * sizeof(float4),
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

int bte_batch_read_framebuffer(int slot, uint32_t* hostRGBA8, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    const BatchScene& b = c.batch[slot];
    const size_t N = (size_t)b.w * b.h;
    if (!hostRGBA8 || count < N) return fail(BTE_ERR_INVALID, "framebuffer copy needs %zu pixels", N);
    if (!c.dBatchFb.p) return fail(BTE_ERR_STATE, "no batch framebuffer (render first)");
    HIPCHK(hipMemcpy(hostRGBA8, c.dBatchFb.as<uint32_t>() + b.pixelBase, N * sizeof(uint32_t),
                     hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_batch_read_accum64(int slot, double* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    const BatchScene& b = c.batch[slot];
    const size_t N = (size_t)b.w * b.h;
    if (!hostRGB || count < 3 * N) return fail(BTE_ERR_INVALID, "accum64 copy needs %zu doubles", 3 * N);
    if (!c.accum64 || !c.dBatchSum64.p) return fail(BTE_ERR_STATE, "binary64 accumulation is off");
    HIPCHK(hipMemcpy(hostRGB, c.dBatchSum64.as<double>() + 3 * (size_t)b.pixelBase, 3 * N * sizeof(double),
                     hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_batch_read_sample_radiance(int slot, float* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    if (c.batchLastFrames <= 0 || !c.dFinal.p) return fail(BTE_ERR_STATE, "no batch pass recorded");
    const BatchScene& b = c.batch[slot];
    const size_t P = (size_t)b.w * b.h * (size_t)c.batchLastFrames;
    if (!hostRGB || count < 3 * P) return fail(BTE_ERR_INVALID, "sample copy needs %zu floats", 3 * P);
    const long long base = batchPathBase(c, c.batchLastFrames, slot);
    std::vector<float> tmp(4 * P);
    HIPCHK(hipMemcpy(tmp.data(), c.dFinal.as<float4>() + base, P * sizeof(float4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < P; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

int bte_batch_read_hit_capture(int slot, int* hostTris, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    if (c.batchLastFrames <= 0 || c.hitCaptureRec <= 0 || !c.dHitCap.p)
        return fail(BTE_ERR_STATE, "no batch hit capture recorded");
    const BatchScene& b = c.batch[slot];
    const size_t rec = (size_t)c.hitCaptureRec;
    const size_t n = (size_t)b.w * b.h * (size_t)c.batchLastFrames * rec;
    if (!hostTris || count < n) return fail(BTE_ERR_INVALID, "hit capture copy needs %zu ints", n);
    const long long base = batchPathBase(c, c.batchLastFrames, slot);
    HIPCHK(hipMemcpy(hostTris, c.dHitCap.as<int>() + (size_t)base * rec, n * sizeof(int), hipMemcpyDeviceToHost));
    return BTE_OK;
}

int bte_batch_denoise(void) {
    ENSURE_DEVICE();
    Context& c = ctx();
    if (c.batch.empty()) return fail(BTE_ERR_STATE, "the batch has no scenes");
    if (!c.aov || !c.dBatchAovAlb.p || c.batchFrame == 0)
        return fail(BTE_ERR_STATE, "batch denoising needs AOVs (bte_enable_aov before bte_batch_render) and a render");
    CHECK_RC(c.dBatchDenoised.ensure((size_t)c.batchPixels * sizeof(float4)));
    for (const BatchScene& b : c.batch) {
        RenderView v;
        v.stream = c.stream;
        v.device = c.device;
        v.width = b.w;
        v.height = b.h;
        v.accum = c.dBatchAccum.as<float4>() + b.pixelBase;
        v.aovAlb = c.dBatchAovAlb.as<float4>() + b.pixelBase;
        v.aovNrm = c.dBatchAovNrm.as<float4>() + b.pixelBase;
        v.fb = c.dBatchFb.as<uint32_t>() + b.pixelBase;
        v.frameIndex = c.batchFrame;
        CHECK_RC(denoise_view(v, c.dBatchDenoised.as<float4>() + b.pixelBase));
    }
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(hipGetLastError());
    c.batchDenoised = true;
    return BTE_OK;
}

int bte_batch_read_denoised(int slot, float* hostRGB, size_t count) {
    ENSURE_DEVICE();
    Context& c = ctx();
    CHECK_RC(batchSlot(c, slot));
    const BatchScene& b = c.batch[slot];
    const size_t N = (size_t)b.w * b.h;
    if (!hostRGB || count < 3 * N) return fail(BTE_ERR_INVALID, "denoised copy needs %zu floats", 3 * N);
    if (!c.batchDenoised) return fail(BTE_ERR_STATE, "the batch has not been denoised since its last render");
    std::vector<float> tmp(4 * N);
    HIPCHK(hipMemcpy(tmp.data(), c.dBatchDenoised.as<float4>() + b.pixelBase, N * sizeof(float4),
                     hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) {
        hostRGB[3 * i] = tmp[4 * i];
        hostRGB[3 * i + 1] = tmp[4 * i + 1];
        hostRGB[3 * i + 2] = tmp[4 * i + 2];
    }
    return BTE_OK;
}

}  // extern "C"
