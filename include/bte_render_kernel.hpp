// bte_render_kernel.hpp — C++ host layer with the reference's API names.
//
// Header-only replacement for include/render_kernel.h of the reference
// (Anil-Parthasarathi/blueTruthEngine): the same free functions, argument
// lists and POD types, forwarded to the C ABI of libbte.so (bte_api.h).  Like
// the reference (src/host/cuda_check.h:14-32) every function fails fast: on a
// non-zero status it prints the library's message and calls std::exit.  Code
// that wants to observe errors calls the bte_* functions directly.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "bte_api.h"

using Float2 = BteFloat2;
using Float3 = BteFloat3;
using Float4 = BteFloat4;
using TriangleData = BteTriangle;
using CameraData = BteCamera;
using EmitterData = BteEmitter;
using SpotlightData = BteSpotlight;
using BsdfData = BteBsdf;

enum BsdfType : int {
    BSDF_Diffuse = BTE_BSDF_DIFFUSE,
    BSDF_Dielectric = BTE_BSDF_DIELECTRIC,
    BSDF_Mirror = BTE_BSDF_MIRROR,
    BSDF_Microfacet = BTE_BSDF_MICROFACET,
    BSDF_Disney = BTE_BSDF_DISNEY,
};

enum class RenderMode { Megakernel, Wavefront };

namespace bte_compat {
inline void check(int rc, const char* fn) {
    if (rc < 0) {
        std::fprintf(stderr, "[bte] %s failed (%d): %s\n", fn, rc, bte_last_error());
        std::exit(EXIT_FAILURE);
    }
}
}  // namespace bte_compat

inline void cudaInitOptix() { bte_compat::check(bte_build_accel(), "cudaInitOptix"); }
inline void cudaInitScene(const TriangleData* triangles, int triangleCount, const Float3* materials, int materialCount,
                          const int* triangleMaterialIds, int imageWidth, int imageHeight) {
    bte_compat::check(bte_init_scene(triangles, triangleCount, materials, materialCount, triangleMaterialIds,
                                     imageWidth, imageHeight), "cudaInitScene");
}
inline void cudaInit(const TriangleData& tri, const Float3& color, int imageWidth, int imageHeight) {
    bte_compat::check(bte_init(&tri, color, imageWidth, imageHeight), "cudaInit");
}
inline void cudaInitCamera(const CameraData& camera) { bte_compat::check(bte_init_camera(&camera), "cudaInitCamera"); }
inline void cudaInitTriangleEmission(const Float3* e, int n) {
    bte_compat::check(bte_init_triangle_emission(e, n), "cudaInitTriangleEmission");
}
inline void cudaInitTriangleEmitterFlags(const uint8_t* f, int n) {
    bte_compat::check(bte_init_triangle_emitter_flags(f, n), "cudaInitTriangleEmitterFlags");
}
inline void cudaInitEmitterTable(const EmitterData* emitters, int emitterCount, const int* emitterTriIndices,
                                 int emitterTriIndexCount, const float* emitterTriCdf, int emitterTriCdfCount,
                                 const float* sceneEmitterCdf) {
    bte_compat::check(bte_init_emitter_table(emitters, emitterCount, emitterTriIndices, emitterTriIndexCount,
                                             emitterTriCdf, emitterTriCdfCount, sceneEmitterCdf),
                      "cudaInitEmitterTable");
}
inline void cudaInitBsdfs(const BsdfData* bsdfs, int bsdfCount, const int* triangleBsdfIds, int triangleCount) {
    bte_compat::check(bte_init_bsdfs(bsdfs, bsdfCount, triangleBsdfIds, triangleCount), "cudaInitBsdfs");
}
inline void cudaInitSpotlights(const SpotlightData* s, int n) {
    bte_compat::check(bte_init_spotlights(s, n), "cudaInitSpotlights");
}
inline void cudaInitTextures(const uint8_t* const* pixels, const int* widths, const int* heights, int count) {
    bte_compat::check(bte_init_textures(pixels, widths, heights, count), "cudaInitTextures");
}
inline void cudaInitMrTextures(const uint8_t* const* pixels, const int* widths, const int* heights, int count) {
    bte_compat::check(bte_init_mr_textures(pixels, widths, heights, count), "cudaInitMrTextures");
}
inline void cudaInitEnvMap(const float* rgba, int width, int height, float intensity, float rotationDeg) {
    bte_compat::check(bte_init_env_map(rgba, width, height, intensity, rotationDeg), "cudaInitEnvMap");
}
// Headless stand-in for the GL PBO: a device buffer of W*H RGBA8 pixels (or 0
// for the library-owned framebuffer).
inline void cudaRegisterPBO(uint32_t* deviceRGBA8) {
    bte_compat::check(bte_register_framebuffer(deviceRGBA8), "cudaRegisterPBO");
}
inline void cudaResetAccumulation(int w, int h) { bte_compat::check(bte_reset_accumulation(w, h), "cudaResetAccumulation"); }
inline void cudaRender(int w, int h) { bte_compat::check(bte_render(w, h), "cudaRender"); }
inline bool cudaToggleDenoiser() {
    const int rc = bte_toggle_denoiser();
    return rc > 0;   // stays off until denoiser weights are loaded
}
inline void cudaSetRenderMode(RenderMode mode) {
    bte_compat::check(bte_set_render_mode(mode == RenderMode::Wavefront ? BTE_MODE_WAVEFRONT : BTE_MODE_MEGAKERNEL),
                      "cudaSetRenderMode");
}
inline RenderMode cudaGetRenderMode() {
    return bte_get_render_mode() == BTE_MODE_WAVEFRONT ? RenderMode::Wavefront : RenderMode::Megakernel;
}
inline void cudaCleanup() { bte_compat::check(bte_cleanup(), "cudaCleanup"); }

// The scene set-up of the reference's main() (src/main.cpp:524-935) in one
// call: loadSceneDescription + OBJ import + BSDF/emitter tables + spotlights +
// .hdr environment, uploaded in the reference's order.  Relative paths resolve
// against the working directory, as in the reference.  Returns the <window>
// size; exits with a message on failure (the reference's behaviour).
struct BteWindowSize {
    int width, height;
};
inline BteWindowSize loadSceneXml(const char* xmlPath) {
    BteScene* scene = nullptr;
    if (bte_scene_load_xml(xmlPath, "", &scene) != BTE_OK) {
        std::fprintf(stderr, "[scene] %s\n", bte_scene_last_error());
        std::exit(EXIT_FAILURE);
    }
    BteSceneInfo info{};
    bte_scene_info(scene, &info);
    const int rc = bte_scene_upload(scene);
    bte_scene_free(scene);
    bte_compat::check(rc, "loadSceneXml");
    return BteWindowSize{info.width, info.height};
}
