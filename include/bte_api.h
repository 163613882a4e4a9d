/* bte_api.h — C ABI of the MI355X wavefront path tracer (libbte.so).
 *
 * Drop-in boundary for the reference's public renderer API
 * (include/render_kernel.h:96-211 of Anil-Parthasarathi/blueTruthEngine).
 * Every entry point keeps the reference's argument list and meaning; the
 * differences are the ones a headless, non-exiting C ABI needs:
 *   - functions return an int status (BTE_OK = 0, negative = error class)
 *     instead of calling exit(); bte_last_error() describes the last failure;
 *   - rendering targets a device framebuffer owned by the library (or one the
 *     caller registers) instead of an OpenGL PBO;
 *   - the OptiX acceleration-structure build becomes bte_build_accel().
 * All scene uploads copy synchronously from caller memory, as the reference
 * does.  The API is single-threaded; calls return after the GPU work they
 * issue has completed.
 */
#ifndef BTE_API_H
#define BTE_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POD types: byte-identical to include/render_kernel.h:9-94 ---------- */
typedef struct { float x, y; } BteFloat2;
typedef struct { float x, y, z; } BteFloat3;
typedef struct { float x, y, z, w; } BteFloat4;

typedef struct {                       /* TriangleData, render_kernel.h:22-26 (96 B) */
    BteFloat3 v0, v1, v2;
    BteFloat3 n0, n1, n2;
    BteFloat2 uv0, uv1, uv2;
} BteTriangle;

typedef struct {                       /* CameraData, render_kernel.h:33-40 (56 B) */
    BteFloat3 origin, forward, right, up;
    float fovYRadians, aspect;
} BteCamera;

typedef struct {                       /* EmitterData, render_kernel.h:44-51 (32 B) */
    int triIndexOffset, triCount, cdfOffset;
    BteFloat3 radiance;
    float areaSum, powerWeight;
} BteEmitter;

typedef struct {                       /* SpotlightData, render_kernel.h:56-63 (48 B) */
    BteFloat3 position, direction, radiance;
    float intensity, innerConeCosine, outerConeCosine;
} BteSpotlight;

typedef struct {                       /* BsdfData, render_kernel.h:88-94 (68 B) */
    int type;                          /* BteBsdfType */
    BteFloat4 p0, p1, p2, p3;
} BteBsdf;

enum BteBsdfType {                     /* BsdfType, render_kernel.h:79-85 */
    BTE_BSDF_DIFFUSE = 1,
    BTE_BSDF_DIELECTRIC = 2,
    BTE_BSDF_MIRROR = 3,
    BTE_BSDF_MICROFACET = 4,
    BTE_BSDF_DISNEY = 5
};

enum BteRenderMode {                   /* RenderMode, render_kernel.h:199-202 */
    BTE_MODE_MEGAKERNEL = 0,
    BTE_MODE_WAVEFRONT = 1
};

enum BteStatus {
    BTE_OK = 0,
    BTE_ERR_INVALID = -1,              /* bad argument / call order */
    BTE_ERR_DEVICE = -2,               /* HIP runtime failure */
    BTE_ERR_MEMORY = -3,               /* device allocation failed */
    BTE_ERR_STATE = -4,                /* required scene data missing */
    BTE_ERR_UNSUPPORTED = -5           /* feature not available in this build */
};

/* Renderer configuration (the reference's compile-time constants,
 * rt_constants.cuh:11-19, made runtime). */
typedef struct {
    int maxDepth;            /* bounce cap, reference MAX_BOUNCES = 200 */
    int megakernelSpp;       /* paths per pixel per frame in megakernel mode (reference 5) */
    int framesInFlight;      /* wavefront: frames traced concurrently (0 = auto) */
    int memoryBudgetMB;      /* cap for wavefront path state (0 = default 32768) */
    int traversalStack;      /* BVH stack entries kept in LDS per lane: 8, 16 or 32 (0 = default 16) */
    int graphs;              /* 1 = replay the bounce loop as captured HIP graphs (default), 0 = plain launches */
    int workChunk;           /* 64-slot queue rows a persistent wave takes per atomic: 1, 2, 4, 8 (0 = 4) */
    int traversalBlocksPerCU;/* resident traversal blocks per CU (0 = as many as occupancy allows) */
    int bvhWidth;            /* 2: BVH2, 4: BVH4 or 8: BVH8 (both collapsed from the BVH2) for extend / connect (default 4; 0 = 2) */
    int tailBounce;          /* wavefront: from this bounce on one fused pass finishes every path
                                (0 = per-bounce passes to the end; default 6) */
} BteConfig;

/* Counters of the last bte_render / bte_render_frames call. */
typedef struct {
    double renderMs;         /* wall time of the GPU work, HIP events */
    long long extensionRays; /* closest-hit rays traced */
    long long hits;          /* extension rays that hit a triangle */
    long long shadowRays;    /* any-hit rays traced */
    long long pathsTraced;   /* primary paths (pixel samples) */
    int bounces;             /* bounce iterations executed */
    int framesInFlight;      /* frames per wavefront pass */
    int bvhNodes;            /* nodes of the current BVH2 */
    int bvhDepth;            /* depth of the current BVH2 */
    long long raysPerBounce[64]; /* extension rays per bounce iteration (first 64) */
    /* filled when bte_set_instrumentation enables them: */
    double extendMs, shadeMs, connectMs;   /* summed kernel time (HIP events) */
    int extendLaunches, connectLaunches;
    long long extendNodes, extendTris;     /* BVH nodes visited / triangles tested */
    long long connectNodes, connectTris;
    int graphLaunches;       /* captured bounce-loop segments replayed by this call */
    double tailMs;           /* kernel time of the tail passes (instrumented; not in extendMs) */
    int tailLaunches;
} BteStats;

enum BteInstrument {
    BTE_INSTR_KERNEL_TIMES = 1,        /* HIP events around every extend/shade/connect launch */
    BTE_INSTR_TRAVERSAL_COUNTS = 2     /* count BVH nodes and triangle tests (slower kernels; the
                                          per-bounce passes then run to the last bounce, no tail pass) */
};

/* ---- scene upload: render_kernel.h:103-173 ------------------------------ */
/* ≙ cudaInitScene (render_kernel.h:111).  materials are stored but, as in the
 * reference (scene_upload.cu:48-54), never read by the integrator. */
int bte_init_scene(const BteTriangle* triangles, int triangleCount,
                   const BteFloat3* materials, int materialCount,
                   const int* triangleMaterialIds, int imageWidth, int imageHeight);
/* ≙ cudaInit (render_kernel.h:117): single-triangle convenience form. */
int bte_init(const BteTriangle* tri, BteFloat3 color, int imageWidth, int imageHeight);
/* ≙ cudaInitCamera (render_kernel.h:121). */
int bte_init_camera(const BteCamera* camera);
/* ≙ cudaInitTriangleEmission (render_kernel.h:125); null/0 clears. */
int bte_init_triangle_emission(const BteFloat3* triangleEmission, int triangleCount);
/* ≙ cudaInitTriangleEmitterFlags (render_kernel.h:129); null/0 clears. */
int bte_init_triangle_emitter_flags(const uint8_t* flags, int triangleCount);
/* ≙ cudaInitEmitterTable (render_kernel.h:137); null/0 clears. */
int bte_init_emitter_table(const BteEmitter* emitters, int emitterCount,
                           const int* emitterTriIndices, int emitterTriIndexCount,
                           const float* emitterTriCdf, int emitterTriCdfCount,
                           const float* sceneEmitterCdf);
/* ≙ cudaInitBsdfs (render_kernel.h:143). */
int bte_init_bsdfs(const BteBsdf* bsdfs, int bsdfCount, const int* triangleBsdfIds, int triangleCount);
/* ≙ cudaInitSpotlights (render_kernel.h:148); any count; null/0 clears. */
int bte_init_spotlights(const BteSpotlight* spotlights, int spotlightCount);
/* ≙ cudaInitTextures / cudaInitMrTextures (render_kernel.h:155,163): entry i
 * (RGBA8, row-major, widths[i] x heights[i]) serves material i of
 * bte_init_scene's triangleMaterialIds; a null entry has none.  Bilinear, wrap
 * addressing, bytes read as byte / 255.  The albedo map replaces BsdfData p0.xyz
 * (microfacet ks recomputed); the metallic-roughness map (G = roughness,
 * B = metallic) scales a Disney BSDF's p0.w and p1.x, for materials that also
 * lie within the albedo set (rt_material.cuh:26-56).  Null/0 clears. */
int bte_init_textures(const uint8_t* const* pixels, const int* widths, const int* heights, int textureCount);
int bte_init_mr_textures(const uint8_t* const* pixels, const int* widths, const int* heights, int textureCount);
/* ≙ cudaInitEnvMap (render_kernel.h:173).  The marginal/conditional CDFs are
 * built on the GPU in binary64 and stored as binary32.  null/0 clears. */
int bte_init_env_map(const float* hdriPixelsRGBA, int width, int height, float intensity, float rotationDeg);
/* ≙ cudaInitOptix (render_kernel.h:103): builds the BVH2 over the uploaded
 * triangles (and the BVH4 / BVH8 collapsed from it).  Rendering builds it
 * implicitly when the triangles changed. */
int bte_build_accel(void);
/* Builder of bte_build_accel (no counterpart in the reference, whose builder
 * is optixAccelBuild on the device): 1 (default) = on the GPU (clustering of
 * Morton-sorted triangles, surface-area leaves and wide collapse); 0 = binned
 * SAH on the host.  Hits do not depend on the tree.  A GPU tree deeper than
 * the traversal stacks allow is built again on the host. */
int bte_set_accel_builder(int builder);
typedef struct {
    double buildMs;            /* wall time of the last build, upload included */
    int builder;               /* 0 host, 1 GPU */
    int triangles;
    int nodes2, nodes4, nodes8;/* nodes8 = 0 when the BVH8 is not kept */
    int depth2, depth4, depth8;
} BteAccelStats;
/* Of the last BVH build (any scene, batch slots included). */
int bte_get_accel_stats(BteAccelStats* stats);

/* ---- framebuffer / accumulation / rendering: render_kernel.h:177-211 ---- */
/* ≙ cudaRegisterPBO (render_kernel.h:177): a caller-owned device buffer of
 * imageWidth*imageHeight RGBA8 pixels (ABGR packed); null = library-owned. */
int bte_register_framebuffer(uint32_t* deviceRGBA8);
/* ≙ cudaResetAccumulation (render_kernel.h:183). */
int bte_reset_accumulation(int imageWidth, int imageHeight);
/* ≙ cudaRender (render_kernel.h:187): one frame (wavefront: 1 sample per
 * pixel; megakernel: megakernelSpp paths per pixel). */
int bte_render(int imageWidth, int imageHeight);
/* nframes consecutive frames in as few wavefront passes as memory allows;
 * identical result to nframes calls of bte_render. */
int bte_render_frames(int imageWidth, int imageHeight, int nframes);
/* ≙ cudaToggleDenoiser (render_kernel.h:192; src/host/denoiser.cu:111-116):
 * flips the denoise-after-render switch and returns the new state (1 on,
 * 0 off).  When on, every bte_render / bte_render_frames call ends with
 * bte_denoise (src/render_kernel.cu:136-140).  Needs loaded weights. */
int bte_toggle_denoiser(void);
/* ≙ cudaSetRenderMode / cudaGetRenderMode (render_kernel.h:205,208). */
int bte_set_render_mode(int mode);
int bte_get_render_mode(void);
/* ≙ cudaCleanup (render_kernel.h:211). */
int bte_cleanup(void);

/* ---- headless extensions ------------------------------------------------ */
int bte_set_config(const BteConfig* config);
int bte_get_config(BteConfig* config);
int bte_get_default_config(BteConfig* config);   /* the values a new context starts with */
int bte_get_stats(BteStats* stats);
int bte_set_instrumentation(int flags);
const char* bte_last_error(void);
/* Current frame index (samples accumulated since the last reset). */
int bte_frame_index(void);
/* Sets the index (and so the seeds) of the next frame without touching the
 * accumulation: frames [k, k + n) give a sample stream independent of frames
 * [0, n).  The binary32 running mean takes frame f with Welford count f + 1,
 * so an offset stream is read from the binary64 sums (bte_read_accum64). */
int bte_set_frame_index(int frameIndex);
/* Copies of the display buffer (W*H uint32) and the binary32 running mean
 * (W*H*3 floats, row-major, row 0 = bottom image row as in the reference). */
int bte_read_framebuffer(uint32_t* hostRGBA8, size_t count);
int bte_read_accum(float* hostRGB, size_t count);
/* Albedo and shading-normal AOVs (no counterpart in the reference; they feed
 * the denoiser): per sample, the base colour (BsdfData p0.xyz) and the
 * interpolated shading normal at the path's first non-delta hit (diffuse,
 * microfacet or Disney), zero when the path ends before one; averaged per
 * pixel with the radiance's Welford update.  Wavefront mode only.  Enabling
 * clears the AOV means; bte_read_aov copies W*H*3 floats per non-null buffer. */
int bte_enable_aov(int enable);
int bte_read_aov(float* hostAlbedo, float* hostNormal, size_t count);
/* Binary64 accumulation of plain sample sums (oracle images). */
int bte_enable_accum64(int enable);
int bte_read_accum64(double* hostRGB, size_t count);
/* Device pointer of the binary32 running mean (float4 per pixel, w unused). */
int bte_accum_device_ptr(void** devicePtr);
/* Per-sample radiance of the last wavefront pass, indexed by
 * (frame - firstFrameOfPass) * W*H + pixel; 3 floats each. */
int bte_read_sample_radiance(float* hostRGB, size_t count);
/* Record the hit triangle of the first maxRecord bounces of every path of the
 * next wavefront pass (-1 miss, -2 not traced); 0 disables. */
int bte_set_hit_capture(int maxRecord);
int bte_read_hit_capture(int* hostTris, size_t count);
/* Env-map CDF arrays as built on the GPU (marginal H+1, conditional H*(W+1)). */
int bte_read_env_cdf(float* marginal, float* conditional, float* totalPower);
/* ---- multi-scene batches ------------------------------------------------ */
/* No counterpart in the reference (its host renders one scene per process,
 * src/main.cpp:846-956): scenes of uneven size and resolution traced together,
 * every bounce of every scene in one persistent wavefront launch.
 * bte_batch_add_scene moves the scene set up through the bte_init_* calls
 * (triangles, BSDFs, emitters, spotlights, env map, camera; the BVH is built
 * here) into a new batch slot rendered at width x height, leaves the current
 * scene empty for the next one, resets every slot's accumulation and returns
 * the slot index (at most 255 slots). */
int bte_batch_add_scene(int width, int height);
int bte_batch_scene_count(void);
/* nframes frames (1 sample per pixel each) of every slot, accumulated per slot
 * exactly as bte_render_frames would accumulate that scene alone (same seeds,
 * same pixels); the current BteConfig (maxDepth, framesInFlight, memory budget,
 * launch shape, graphs) applies to the whole batch.  Wavefront mode only. */
int bte_batch_render(int nframes);
int bte_batch_reset(void);          /* zero every slot's accumulation; frame index 0 */
int bte_batch_clear(void);          /* free every slot (bte_cleanup does too) */
int bte_batch_frame_index(void);
/* Per slot, as bte_read_accum / bte_read_framebuffer / bte_read_accum64. */
int bte_batch_read_accum(int slot, float* hostRGB, size_t count);
int bte_batch_read_framebuffer(int slot, uint32_t* hostRGBA8, size_t count);
int bte_batch_read_accum64(int slot, double* hostRGB, size_t count);
/* Per slot, as bte_read_sample_radiance / bte_read_hit_capture, for the
 * batch's last wavefront pass: (frame - firstFrameOfPass) * W*H + pixel of
 * that slot's scene (bte_set_hit_capture applies to batches too). */
int bte_batch_read_sample_radiance(int slot, float* hostRGB, size_t count);
int bte_batch_read_hit_capture(int slot, int* hostTris, size_t count);
/* Per slot, as bte_read_env_cdf. */
int bte_batch_read_env_cdf(int slot, float* marginal, float* conditional, float* totalPower);
/* The U-Net over every slot (AOVs enabled before the batch's first render),
 * into each slot's display buffer and its denoised linear image. */
int bte_batch_denoise(void);
int bte_batch_read_denoised(int slot, float* hostRGB, size_t count);
/* ---- learned denoiser (the slot of src/host/denoiser.cu:72-109) --------- */
/* ≙ denoiseAndPresent (denoiser.cu:72-109): a U-Net (4 levels, 32/64/128/128
 * channels, bf16 MFMA convolutions) over log(1+RGB) of the accumulation and
 * the albedo / normal AOVs (bte_enable_aov) writes the denoised linear image
 * (bte_read_denoised) and its gamma-packed pixels into the display buffer. */
int bte_denoise(void);
int bte_read_denoised(float* hostRGB, size_t count);
/* Parameters in canonical order, layer by layer (c0a c0b c1a c1b c2a c2b c3a
 * c3b d2a d2b d1a d1b d0a d0b out): W[cout][3][3][cin] then b[cout], fp32. */
long long bte_denoiser_param_count(void);
int bte_denoiser_init(unsigned seed);                 /* He-normal weights, zero biases */
/* bte_cleanup switches the denoiser off and frees its per-resolution buffers
 * but keeps the weights and training pairs; this frees those too. */
int bte_denoiser_release(void);
int bte_denoiser_set_weights(const float* params, long long count);
/* The same from a NumPy .npy file of little-endian float32 (e.g. the shipped
 * bluetruthengine_amd/weights/unet_c4.npy). */
int bte_denoiser_load_file(const char* path);
int bte_denoiser_get_weights(float* params, long long count);
int bte_denoiser_get_grads(float* grads, long long count);   /* of the last training step */
/* Training on this renderer's own pairs: reserve slots of width x height
 * (width a multiple of 256, height of 8), capture the current accumulation as
 * a slot's input (which = 0: features incl. AOVs, e.g. after 4 spp) or target
 * (which = 1: log(1+RGB), e.g. after 4096 spp), then run AdamW steps on random
 * batches of complete pairs; losses (optional) receives each step's mean L1. */
/* Training loss: 0 = L1 of log(1 + RGB) (default), 1 = relative MSE of the
 * linear radiance, (y - t)^2 / (y^2 + 0.01) with the prediction y held
 * constant in the denominator (the relMSE used to judge images). */
int bte_denoiser_set_loss(int mode);
/* Inference flip ensemble (no counterpart in the reference, whose OptiX
 * denoiser is a single pass): bte_denoise / cudaToggleDenoiser run the network
 * on `flips` copies of the frame (1: as is; 2: and mirrored left-right; 4: and
 * mirrored top-bottom, both ways) as one batch and average the outputs in log
 * space.  Default 4; the time of one denoise grows about in proportion. */
int bte_denoiser_set_flips(int flips);
/* Convolution kernel of the denoiser (no counterpart in the reference; for
 * tests and measurements): 0 = fragments from L1/L2, 1 = weight tiles staged
 * in LDS per K step, 2 (default) = input tile with halo and weight slice
 * staged in LDS per 32-channel slice.  All three compute the same
 * convolution; sums may differ in the last fp32 bits. */
int bte_denoiser_set_conv_kernel(int variant);
/* ≙ optixDenoiserInvoke on caller-given layers (src/host/denoiser.cu:72-102):
 * denoises device buffers of width x height float4 (RGB + unused w; row 0 =
 * bottom, as bte_read_accum) -- radiance, albedo and normal -- into `out`
 * (float4, the same layout), synchronously on the library's stream.  Needs
 * loaded weights; the buffers must be 16-byte aligned. */
int bte_denoise_image(const float* rgb, const float* albedo, const float* normal, int width, int height,
                      float* out);
int bte_denoiser_get_flips(void);
int bte_denoiser_dataset_reserve(int slots, int width, int height);
int bte_denoiser_capture(int slot, int which);
int bte_denoiser_train(int steps, int batch, float lr, float weightDecay, unsigned seed, float* losses);
/* Host-array forms (tests): x9 is n*h*w*9 floats (log(1+RGB), albedo,
 * normal; rounded to bf16 on upload), out3 / target3 are n*h*w*3 log-space
 * values; h, w multiples of 8 (training: w a multiple of 256).  lr = 0 in a
 * training step computes loss and gradients without updating. */
int bte_denoiser_forward_host(const float* x9, int n, int h, int w, float* out3);
int bte_denoiser_train_step_host(const float* x9, const float* target3, int n, int h, int w, float lr,
                                 float weightDecay, float* loss);
/* Intermediate tensors of the last pass as fp32 (NHWC); id 18 + k are the
 * gradients, id 39 the fp32 output; shape4 = {n, h, w, c}. */
int bte_denoiser_tensor_shape(int id, int* shape4);
int bte_denoiser_read_tensor(int id, float* out, long long count);

/* ---- scene host layer: scene.xml / OBJ / .hdr (src/scene.cpp, src/main.cpp) */
/* ≙ loadSceneDescription (src/scene.cpp:83-523) plus the set-up of
 * src/main.cpp:524-930: parses a scene.xml, imports its OBJ meshes (normal
 * and planar-UV fallbacks, Scale->Rx->Ry->Rz->Translate transforms) and its
 * glTF / GLB models (loadGltfModel, src/gltf_loader.cpp:768-846: node graph,
 * Disney materials, albedo and metallic-roughness maps, emissive materials,
 * backdrop / outline / outlier pruning), decodes the materials' albedo
 * textures (PNG or baseline JPEG; a missing file means no texture, as in the
 * reference; an undecodable one is an error), packs the BSDF table, builds
 * the emitter tables, the camera basis, spotlights and the Radiance .hdr
 * environment.  Relative paths resolve against baseDir (the
 * reference resolves them against its working directory).  Host only: no
 * device needed until bte_scene_upload.  Errors: bte_scene_last_error(). */
typedef struct BteScene BteScene;      /* opaque */
typedef struct {
    int width, height;                 /* <window> */
    int triangleCount, materialCount, bsdfCount, emitterCount, emitterTriCount, emitterCdfCount;
    int spotlightCount;
    int envWidth, envHeight;           /* 0 when the scene has no <environment> */
    float envIntensity, envRotationDeg;
    BteCamera camera;
} BteSceneInfo;
typedef struct {                       /* views into the scene's arrays (valid until bte_scene_free) */
    const BteTriangle* triangles;
    const BteFloat3* materials;
    const int* triangleMaterialIds;
    const int* triangleBsdfIds;
    const BteBsdf* bsdfs;
    const BteFloat3* triangleEmission;
    const uint8_t* triangleEmitterFlags;
    const BteEmitter* emitters;
    const int* emitterTriIndices;
    const float* emitterTriCdf;
    const float* sceneEmitterCdf;      /* emitterCount + 1 entries */
    const BteSpotlight* spotlights;
    const float* envRGBA;              /* envWidth * envHeight * 4, row 0 = top; null without env */
    /* albedo textures, materialCount entries (src/main.cpp:600-612): RGBA8,
     * row 0 = top; a null entry = no texture (width = height = 1) */
    const uint8_t* const* albedoTextures;
    const int* albedoTextureWidths;
    const int* albedoTextureHeights;
    /* metallic-roughness maps, materialCount entries (glTF materials only,
     * src/main.cpp:776-781): G = roughness, B = metallic */
    const uint8_t* const* mrTextures;
    const int* mrTextureWidths;
    const int* mrTextureHeights;
} BteSceneArrays;
int bte_scene_load_xml(const char* path, const char* baseDir, BteScene** scene);
const char* bte_scene_last_error(void);
int bte_scene_info(const BteScene* scene, BteSceneInfo* info);
int bte_scene_arrays(const BteScene* scene, BteSceneArrays* arrays);
/* bte_init_scene … bte_init_env_map and bte_init_camera in the reference's order */
int bte_scene_upload(const BteScene* scene);
void bte_scene_free(BteScene* scene);

/* Launch-shape autotuner (no counterpart in the reference, whose launch
 * shape OptiX fixes).  Renders `frames` frames at width x height once per
 * candidate (BVH width 2/4/8, then traversal LDS stack 8/16/32, then work chunk
 * 1/2/4/8 rows, then traversal blocks per CU: all, 3/4, 1/2), by coordinate
 * descent from the current configuration, keeps the fastest in the
 * configuration and resets the accumulation.  Images do not depend on these settings.  msBest
 * (optional) receives the best render time. */
int bte_autotune(int width, int height, int frames, double* msBest);
/* The same search over bte_batch_render(frames) of the current batch (whose
 * accumulation it resets). */
int bte_batch_autotune(int frames, double* msBest);

/* Trace caller rays with the BVH2 kernels: rays = n x (ox,oy,oz,dx,dy,dz).
 * anyHit = 0: closest hit in [tmin, tmax] -> tri (-1 miss), t, u, v.
 * anyHit = 1: occlusion in [tmin, tmaxPerRay[i]] -> tri = 1 occluded / 0. */
int bte_trace_rays(const float* rays, const float* tmaxPerRay, long long n, float tmin, float tmax,
                   int anyHit, int* tri, float* t, float* u, float* v);

/* No counterpart in the reference: bte_trace_rays for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every buffer is device memory of the
 * current device, owned by the caller: rays (n x 6 floats), tmaxPerRay
 * (optional, any hit), tri and, for closest hits, t, u, v.  It traces the
 * current scene's BVH, which bte_build_accel must already have built (it
 * builds, uploads and allocates nothing: no hipMalloc, hipFree or allocating
 * library call).  Its work counter and traversal stacks live in one static
 * device array of 16 MiB + 256 B that each GPU the library runs on holds
 * until the process ends; calls that overlap on one GPU race on it.  Limits:
 * n <= 2^31 - 2^20 rays per call and a BVH of at most 64 levels (the builder
 * guarantees 60); beyond them it returns an error.  Returns after the work. */
int bte_trace_rays_noalloc(const float* rays, const float* tmaxPerRay, long long n, float tmin, float tmax,
                           int anyHit, int* tri, float* t, float* u, float* v);

#ifdef __cplusplus
}
#endif

#endif /* BTE_API_H */
