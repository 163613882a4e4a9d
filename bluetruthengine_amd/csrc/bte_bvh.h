// bte_bvh.h — host BVH2 builder interface.
#pragma once

#include <vector>

#include "bte_types.h"

namespace bte {

struct BvhBuild {
    std::vector<BvhHostNode> nodes;   // node 0 is the (inner) root
    std::vector<float> leafTris;      // 12 floats per leaf-ordered triangle
    int depth = 0;
};

BvhBuild build_bvh2(const BteTriangle* tris, int n);

}  // namespace bte
