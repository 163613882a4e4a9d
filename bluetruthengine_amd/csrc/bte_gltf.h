// bte_gltf.h — glTF 2.0 / GLB models for the scene host layer (<model>).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bte_api.h"

namespace bte {

struct GltfImage {
    std::vector<uint8_t> px;   // RGBA8, row 0 = top; empty = no image
    int width = 0, height = 0;
};

// One material of the model mapped to a Disney BSDF (gltf_loader.h:15-21).
struct GltfSlot {
    BteBsdf bsdf;
    GltfImage albedo, metallicRoughness;
    BteFloat3 emission = {0.0f, 0.0f, 0.0f};
    bool isEmitter = false;
};

struct GltfModel {
    std::vector<BteTriangle> triangles;
    std::vector<int> materialIds;   // per triangle, index into materials
    std::vector<GltfSlot> materials;
};

// The <model> element's placement, applied after the node graph as
// translate * rotZ * rotY * rotX * scale (degrees).
struct ModelPlacement {
    float pos[3] = {0.0f, 0.0f, 0.0f};
    float rotDeg[3] = {0.0f, 0.0f, 0.0f};   // x, y, z
    float scale[3] = {1.0f, 1.0f, 1.0f};
};

// ≙ loadGltfModel (src/gltf_loader.cpp:768-846): a .glb (binary container) or
// .gltf (JSON; buffers and images from data: URIs or files beside it) file
// flattened into world-space triangles and Disney slots.  Returns false with a
// message where the reference exits (unreadable or invalid file, an image
// that does not decode, no triangles).
bool load_gltf(const std::string& path, const ModelPlacement& placement, GltfModel& out, std::string& err);

}  // namespace bte
