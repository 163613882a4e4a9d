// bte_types.h — host/device shared layouts of the wavefront path tracer.
#pragma once

#include <cstdint>

#include "../../include/bte_api.h"

namespace bte {

// Device BVH2 node, 64 B = 4 x float4 (two child boxes per node):
//   n0 = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   n1 = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   n2 = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   n3 = (child0, child1, 0, 0) as int bits
// child >= 0: inner node index; child < 0: leaf, ~child = start << 3 | (count - 1)
// where [start, start+count) indexes the leaf-ordered triangle array.
constexpr int kLeafCountBits = 3;
constexpr int kMaxLeafTris = 1 << kLeafCountBits;
constexpr int kMaxBvhDepth = 60;      // builder guarantee
// Traversal stack entries: a BVH2 visit pushes at most 1 entry per level, a
// BVH4 visit at most 3, and BVH4 depth <= BVH2 depth.
constexpr int kStackSize = 3 * kMaxBvhDepth + 4;

struct BvhHostNode {
    float n[12];
    int child[2];
    int pad[2];
};
static_assert(sizeof(BvhHostNode) == 64, "BVH node must be 64 B");

// Device BVH4 node, 128 B = 8 x float4, child boxes in SoA:
//   q0 = lo.x of children 0..3, q1 = hi.x, q2 = lo.y, q3 = hi.y, q4 = lo.z, q5 = hi.z,
//   q6 = child refs (as in BVH2; empty slots have a far-away point box), q7 = 0.
// Collapsed from the SAH BVH2 (same leaves and triangle order).
struct Bvh4HostNode {
    float box[6][4];
    int child[4];
    int pad[4];
};
static_assert(sizeof(Bvh4HostNode) == 128, "BVH4 node must be 128 B");

// BVH8 node: box[axis lo/hi][child] as in Bvh4HostNode, 8 children, padded to
// two 128-B lines.
struct Bvh8HostNode {
    float box[6][8];
    int child[8];
    int pad[8];
};
static_assert(sizeof(Bvh8HostNode) == 256, "BVH8 node must be 256 B");

// Shadow-record kinds (light classes), traced in this order so contributions
// add in the reference's order: area lights, spotlights (one record per path;
// its pass traces every spotlight in index order), environment.
constexpr int kMaxShadowKinds = 3;

}  // namespace bte
