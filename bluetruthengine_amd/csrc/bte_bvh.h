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

// Collapses a BVH2 into a BVH4 by surface-area cost: each BVH4 node takes
// up to four subtrees of one BVH2 inner node, chosen by dynamic programming
// over the BVH2 (leaves and triangle order are kept).  Node 0 is the root;
// depth4 receives the BVH4 depth (<= BVH2 depth).
std::vector<Bvh4HostNode> collapse_bvh4(const std::vector<BvhHostNode>& nodes2, int& depth4);

// The same collapse into 8-wide nodes (up to eight subtrees per node).  A
// traversal stack needs up to 7 entries per level: (7 * depth8 + 1) entries.
std::vector<Bvh8HostNode> collapse_bvh8(const std::vector<BvhHostNode>& nodes2, int& depth8);

}  // namespace bte
