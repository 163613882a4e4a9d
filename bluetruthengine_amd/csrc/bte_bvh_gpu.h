// bte_bvh_gpu.h — BVH built on the GPU (bte_bvh_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "bte_internal.h"
#include "bte_types.h"

namespace bte {

// Device-resident result: BVH2 / BVH4 / BVH8 nodes and leaf-ordered triangles
// in the layouts of bte_types.h (node 0 the root; every hit-equivalent to the
// host builder's, the trees differ).
struct GpuBvh {
    DBuf dNodes2, dNodes4, dNodes8, dLeafTris;
    int nodes2 = 0, nodes4 = 0, nodes8 = 0;
    int depth2 = 0, depth4 = 0, depth8 = 0;   // root = 1; BVH2 depth counts leaves
    int rounds = 0;                           // clustering rounds
    double buildMs = 0.0;
};

// Builds from n >= 2 triangles already on the device; synchronous on `s`.
int build_bvh_gpu(const BteTriangle* dTris, int n, hipStream_t s, GpuBvh& out);

}  // namespace bte
