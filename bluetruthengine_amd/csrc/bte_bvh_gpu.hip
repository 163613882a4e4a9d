// bte_bvh_gpu.hip — BVH built on the GPU (replaces optixAccelBuild,
// src/host/optix_setup.cu:215-292, on the device as the reference does).
//
// Parallel locally-ordered clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 2018): triangles sorted by the 63-bit Morton code of their box
// centres, then rounds in which every cluster finds its nearest neighbour
// (smallest union-box area) within kRadius positions, mutual pairs merge into
// a new node and the survivors are compacted, until one cluster is left.
// Nodes of round k depend only on earlier rounds, so every later pass walks
// the rounds bottom-up or top-down with one launch per round:
//   bottom-up  subtree triangle counts; leaves of up to kMaxLeafTris
//              triangles where the surface-area cost of a leaf is not above
//              that of the subtree; the surface-area-cost dynamic programs of
//              the 4- and 8-wide collapse (the same costs as bte_bvh.cpp);
//   top-down   leaf ranges in the leaf-ordered triangle array, depths, and the
//              children each wide node collects;
//   emission   exclusive scans number the nodes (root 0, parents before
//              children) and one kernel writes the BVH2, BVH4 and BVH8 nodes
//              and the leaf-ordered triangles in the layouts of bte_types.h.
// Everything is deterministic (stable radix sort, scans, index-ordered ties),
// and hits do not depend on the tree (hitTriangle), only traversal cost does.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <vector>

#include "bte_bvh_gpu.h"
#include "bte_internal.h"

namespace bte {

namespace {

constexpr int kB = 256;           // threads per block
#ifndef BTE_PLOC_RADIUS
#define BTE_PLOC_RADIUS 512
#endif
constexpr int kRadius = BTE_PLOC_RADIUS;   // nearest-neighbour search radius of the clustering
constexpr float kNodeCost = 1.0f; // surface-area costs (as bte_bvh.cpp): a node visit ...
#ifndef BTE_LEAF_TRI_COST
#define BTE_LEAF_TRI_COST 1.5f
#endif
constexpr float kTriCost = BTE_LEAF_TRI_COST;  // ... and a triangle test, per unit of box area (leaf decision)
constexpr float kWideTriCost = 0.3f;   // triangle cost of the wide collapse (bte_bvh.cpp SahCollapser)
constexpr float kFar = -3.0e38f;  // box of an empty child slot: a far-away point

int blocks(long long n) { return (int)((n + kB - 1) / kB); }

__device__ __forceinline__ float boxArea(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
__device__ __forceinline__ float4 vmin(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ __forceinline__ float4 vmax(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

// Spreads the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Triangle boxes (node slots 0..n-1 in input order for now) and centroid bounds.
__global__ void k_tri_boxes(const BteTriangle* tris, int n, float4* lo, float4* hi, float4* cen, float4* blockLo,
                            float4* blockHi) {
    __shared__ float4 sLo[kB], sHi[kB];
    const int i = blockIdx.x * kB + threadIdx.x;
    float4 cl = make_float4(INFINITY, INFINITY, INFINITY, 0), ch = make_float4(-INFINITY, -INFINITY, -INFINITY, 0);
    if (i < n) {
        const BteTriangle t = tris[i];
        const float4 a = make_float4(t.v0.x, t.v0.y, t.v0.z, 0), b = make_float4(t.v1.x, t.v1.y, t.v1.z, 0),
                     c = make_float4(t.v2.x, t.v2.y, t.v2.z, 0);
        const float4 l = vmin(vmin(a, b), c), h = vmax(vmax(a, b), c);
        lo[i] = l;
        hi[i] = h;
        const float4 m = make_float4((l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f, 0.0f);
        cen[i] = m;
        cl = m;
        ch = m;
    }
    sLo[threadIdx.x] = cl;
    sHi[threadIdx.x] = ch;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sLo[threadIdx.x] = vmin(sLo[threadIdx.x], sLo[threadIdx.x + s]);
            sHi[threadIdx.x] = vmax(sHi[threadIdx.x], sHi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        blockLo[blockIdx.x] = sLo[0];
        blockHi[blockIdx.x] = sHi[0];
    }
}

__global__ void k_reduce_bounds(float4* blockLo, float4* blockHi, int nb) {
    __shared__ float4 sLo[kB], sHi[kB];
    float4 l = make_float4(INFINITY, INFINITY, INFINITY, 0), h = make_float4(-INFINITY, -INFINITY, -INFINITY, 0);
    for (int i = threadIdx.x; i < nb; i += kB) {
        l = vmin(l, blockLo[i]);
        h = vmax(h, blockHi[i]);
    }
    sLo[threadIdx.x] = l;
    sHi[threadIdx.x] = h;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sLo[threadIdx.x] = vmin(sLo[threadIdx.x], sLo[threadIdx.x + s]);
            sHi[threadIdx.x] = vmax(sHi[threadIdx.x], sHi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        blockLo[0] = sLo[0];
        blockHi[0] = sHi[0];
    }
}

__global__ void k_morton(const float4* cen, int n, const float4* bounds, uint64_t* keys, int* ids) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const float4 l = bounds[0], h = bounds[kB];   // bounds[kB] holds the high corner (see launcher)
    const float4 c = cen[i];
    const float sx = h.x > l.x ? 2097151.0f / (h.x - l.x) : 0.0f;
    const float sy = h.y > l.y ? 2097151.0f / (h.y - l.y) : 0.0f;
    const float sz = h.z > l.z ? 2097151.0f / (h.z - l.z) : 0.0f;
    const uint64_t x = (uint64_t)fminf(fmaxf((c.x - l.x) * sx, 0.0f), 2097151.0f);
    const uint64_t y = (uint64_t)fminf(fmaxf((c.y - l.y) * sy, 0.0f), 2097151.0f);
    const uint64_t z = (uint64_t)fminf(fmaxf((c.z - l.z) * sz, 0.0f), 2097151.0f);
    keys[i] = spread3(x) << 2 | spread3(y) << 1 | spread3(z);
    ids[i] = i;
}

// Leaf node j (sorted position) takes the box of triangle sortedIds[j].
__global__ void k_leaf_nodes(const int* sortedIds, int n, const float4* triLo, const float4* triHi, float4* nodeLo,
                             float4* nodeHi, int* clusters) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int t = sortedIds[j];
    nodeLo[j] = triLo[t];
    nodeHi[j] = triHi[t];
    clusters[j] = j;
}

// Nearest neighbour of every cluster within kRadius positions: smallest
// union-box area, ties to the smaller (min index, max index) pair, so the
// globally smallest pair is always mutual.  A block stages the boxes of its
// 256 clusters and the kRadius on either side in LDS once.
__global__ __launch_bounds__(kB) void k_nearest(const int* clusters, int m, const float4* nodeLo,
                                                 const float4* nodeHi, int* nn) {
    __shared__ float4 sLo[kB + 2 * kRadius], sHi[kB + 2 * kRadius];
    const int base = (int)blockIdx.x * kB - kRadius;
    for (int t = threadIdx.x; t < kB + 2 * kRadius; t += kB) {
        const int j = base + t;
        if (j >= 0 && j < m) {
            const int c = clusters[j];
            sLo[t] = nodeLo[c];
            sHi[t] = nodeHi[c];
        }
    }
    __syncthreads();
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const float4 li = sLo[i - base], hi = sHi[i - base];
    float best = INFINITY;
    int bj = -1;
    const int j0 = max(0, i - kRadius), j1 = min(m - 1, i + kRadius);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float d = boxArea(vmin(li, sLo[j - base]), vmax(hi, sHi[j - base]));
        // ties: pair keys (min(i, j), max(i, j)) grow with j on each side of i, so the first j wins
        if (d < best) {
            best = d;
            bj = j;
        }
    }
    nn[i] = bj;
}

__global__ void k_merge_flags(const int* nn, int m, int* mergeFlag, int* keepFlag) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    mergeFlag[i] = (mutual && i < j) ? 1 : 0;
    keepFlag[i] = (mutual && i > j) ? 0 : 1;
}

__global__ void k_merge_apply(const int* clusters, const int* nn, int m, const int* mergeFlag, const int* mergeIdx,
                              const int* keepFlag, const int* keepIdx, int nodeBase, int* child0, int* child1,
                              float4* nodeLo, float4* nodeHi, int* next) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    if (mergeFlag[i]) {
        const int node = nodeBase + mergeIdx[i];
        const int a = clusters[i], b = clusters[nn[i]];
        child0[node] = a;
        child1[node] = b;
        nodeLo[node] = vmin(nodeLo[a], nodeLo[b]);
        nodeHi[node] = vmax(nodeHi[a], nodeHi[b]);
        next[keepIdx[i]] = node;
    } else if (keepFlag[i]) {
        next[keepIdx[i]] = clusters[i];
    }
}

// Wide-collapse cost of child c given `slots` (a leaf: its triangles).
template <int S>
__device__ __forceinline__ float wideChildCost(int c, int slots, int n, const uint8_t* leafMark, const int* count,
                                               const float4* nodeLo, const float4* nodeHi, const float* cost) {
    if (c < n || leafMark[c]) return boxArea(nodeLo[c], nodeHi[c]) * kWideTriCost * (float)(c < n ? 1 : count[c]);
    return cost[(size_t)c * S + slots];
}

template <int W>
__device__ void wideSolve(int i, int n, const int* child0, const int* child1, const uint8_t* leafMark,
                          const int* count, const float4* nodeLo, const float4* nodeHi, float* cost, int8_t* split) {
    constexpr int S = W + 1;
    const int a = child0[i], b = child1[i];
    auto distribute = [&](int slots, int& bestK) {
        float best = INFINITY;
        bestK = 1;
        for (int k = 1; k < slots; ++k) {
            const float c = wideChildCost<S>(a, k, n, leafMark, count, nodeLo, nodeHi, cost) +
                            wideChildCost<S>(b, slots - k, n, leafMark, count, nodeLo, nodeHi, cost);
            if (c < best) {
                best = c;
                bestK = k;
            }
        }
        return best;
    };
    int kw = 1;
    const float asNode = boxArea(nodeLo[i], nodeHi[i]) * kNodeCost + distribute(W, kw);
    cost[(size_t)i * S + 1] = asNode;
    split[(size_t)i * S + 0] = (int8_t)kw;
    for (int slots = 2; slots <= W; ++slots) {
        int k = 1;
        const float spread = distribute(slots, k);
        cost[(size_t)i * S + slots] = fminf(asNode, spread);
        split[(size_t)i * S + slots] = spread < asNode ? (int8_t)k : (int8_t)0;
    }
}

// Bottom-up over the nodes of one clustering round.
__global__ void k_bottom_up(int first, int cnt, int n, int root, const int* child0, const int* child1,
                            const float4* nodeLo, const float4* nodeHi, int* count, float* cost2, uint8_t* leafMark,
                            float* cost4, int8_t* split4, float* cost8, int8_t* split8) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= cnt) return;
    const int i = first + k;
    const int a = child0[i], b = child1[i];
    const int na = a < n ? 1 : count[a], nb = b < n ? 1 : count[b];
    const int c = na + nb;
    count[i] = c;
    const float area = boxArea(nodeLo[i], nodeHi[i]);
    const float ca = a < n ? boxArea(nodeLo[a], nodeHi[a]) * kTriCost : cost2[a];
    const float cb = b < n ? boxArea(nodeLo[b], nodeHi[b]) * kTriCost : cost2[b];
    const float asSplit = area * kNodeCost + ca + cb;
    const float asLeaf = area * kTriCost * (float)c;
    const bool leaf = i != root && c <= kMaxLeafTris && asLeaf <= asSplit;
    leafMark[i] = leaf ? 1 : 0;
    cost2[i] = leaf ? asLeaf : asSplit;
    if (!leaf) {
        wideSolve<4>(i, n, child0, child1, leafMark, count, nodeLo, nodeHi, cost4, split4);
        wideSolve<8>(i, n, child0, child1, leafMark, count, nodeLo, nodeHi, cost8, split8);
    }
}

// Children a wide node rooted at BVH2 node i collects (the split decisions of
// the dynamic program, depth first: the slot order of bte_bvh.cpp).
template <int W>
__device__ int wideCollect(int i, int n, const int* child0, const int* child1, const uint8_t* leafMark,
                           const int8_t* split, int* out) {
    constexpr int S = W + 1;
    int stackNode[2 * W], stackSlots[2 * W];
    int sp = 0, cnt = 0;
    const int k0 = split[(size_t)i * S + 0];
    stackNode[sp] = child1[i];
    stackSlots[sp++] = W - k0;
    stackNode[sp] = child0[i];
    stackSlots[sp++] = k0;
    while (sp > 0) {
        --sp;
        const int c = stackNode[sp], slots = stackSlots[sp];
        const int k = (c < n || leafMark[c] || slots == 1) ? 0 : split[(size_t)c * S + slots];
        if (k == 0) {
            out[cnt++] = c;
            continue;
        }
        stackNode[sp] = child1[c];
        stackSlots[sp++] = slots - k;
        stackNode[sp] = child0[c];
        stackSlots[sp++] = k;
    }
    return cnt;
}

// Top-down over the nodes of one clustering round (rounds in reverse order).
// state bits: 1 kept BVH2 inner node, 2 BVH4 node, 4 BVH8 node.
__global__ void k_top_down(int first, int cnt, int n, const int* child0, const int* child1, const uint8_t* leafMark,
                           const int* count, const int8_t* split4, const int8_t* split8, int* offset, int* depth2,
                           int* depth4, int* depth8, uint8_t* state, const int* sortedIds, int* order, int* col4,
                           int* col8, int* maxDepths) {
    const int k = blockIdx.x * kB + threadIdx.x;
    if (k >= cnt) return;
    const int i = first + k;
    const uint8_t st = state[i];
    if (!(st & 1)) {
        if (leafMark[i]) {
            // a BVH2 leaf reached from a kept inner node: its triangles, left to right
            if (offset[i] < 0) return;   // below another leaf (never: leaf subtrees are not visited)
            int stack[kMaxLeafTris], sp = 0, at = offset[i];
            stack[sp++] = i;
            while (sp > 0) {
                const int c = stack[--sp];
                if (c < n) {
                    order[at++] = sortedIds[c];
                } else {
                    stack[sp++] = child1[c];
                    stack[sp++] = child0[c];
                }
            }
        }
        return;
    }
    const int a = child0[i], b = child1[i];
    offset[a] = offset[i];
    offset[b] = offset[i] + (a < n ? 1 : count[a]);
    const int d = depth2[i] + 1;
    depth2[a] = d;
    depth2[b] = d;
    atomicMax(maxDepths + 0, d);
    if (a < n) order[offset[a]] = sortedIds[a];
    if (b < n) order[offset[b]] = sortedIds[b];
    if (a >= n && !leafMark[a]) state[a] |= 1;
    if (b >= n && !leafMark[b]) state[b] |= 1;
    if (st & 2) {
        int* out = col4 + (size_t)i * 4;
        const int c = wideCollect<4>(i, n, child0, child1, leafMark, split4, out);
        for (int s = c; s < 4; ++s) out[s] = -1;
        atomicMax(maxDepths + 1, depth4[i]);
        for (int s = 0; s < c; ++s)
            if (out[s] >= n && !leafMark[out[s]]) {
                state[out[s]] |= 2;
                depth4[out[s]] = depth4[i] + 1;
            }
    }
    if (st & 4) {
        int* out = col8 + (size_t)i * 8;
        const int c = wideCollect<8>(i, n, child0, child1, leafMark, split8, out);
        for (int s = c; s < 8; ++s) out[s] = -1;
        atomicMax(maxDepths + 2, depth8[i]);
        for (int s = 0; s < c; ++s)
            if (out[s] >= n && !leafMark[out[s]]) {
                state[out[s]] |= 4;
                depth8[out[s]] = depth8[i] + 1;
            }
    }
}

// Node numbering flags in descending node order (root first).
__global__ void k_number_flags(int n, int inner, const uint8_t* state, int* f2, int* f4, int* f8) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= inner) return;
    const int i = 2 * n - 2 - j;
    const uint8_t st = state[i];
    f2[j] = (st & 1) ? 1 : 0;
    f4[j] = (st & 2) ? 1 : 0;
    f8[j] = (st & 4) ? 1 : 0;
}

__device__ __forceinline__ int childRef(int c, int n, const uint8_t* leafMark, const int* count, const int* offset,
                                        const int* idx) {
    if (c < n || leafMark[c]) {
        const int cnt = c < n ? 1 : count[c];
        return ~((offset[c] << kLeafCountBits) | (cnt - 1));
    }
    return idx[2 * n - 2 - c];
}

__global__ void k_emit_nodes(int n, int inner, const int* child0, const int* child1, const uint8_t* leafMark,
                             const int* count, const int* offset, const uint8_t* state, const float4* nodeLo,
                             const float4* nodeHi, const int* col4, const int* col8, const int* idx2, const int* idx4,
                             const int* idx8, BvhHostNode* out2, Bvh4HostNode* out4, Bvh8HostNode* out8) {
    const int j = blockIdx.x * kB + threadIdx.x;
    if (j >= inner) return;
    const int i = 2 * n - 2 - j;
    const uint8_t st = state[i];
    if (st & 1) {
        const int a = child0[i], b = child1[i];
        const float4 la = nodeLo[a], ha = nodeHi[a], lb = nodeLo[b], hb = nodeHi[b];
        BvhHostNode nd;
        nd.n[0] = la.x; nd.n[1] = ha.x; nd.n[2] = la.y; nd.n[3] = ha.y;
        nd.n[4] = lb.x; nd.n[5] = hb.x; nd.n[6] = lb.y; nd.n[7] = hb.y;
        nd.n[8] = la.z; nd.n[9] = ha.z; nd.n[10] = lb.z; nd.n[11] = hb.z;
        nd.child[0] = childRef(a, n, leafMark, count, offset, idx2);
        nd.child[1] = childRef(b, n, leafMark, count, offset, idx2);
        nd.pad[0] = nd.pad[1] = 0;
        out2[idx2[j]] = nd;
    }
    if (st & 2) {
        Bvh4HostNode nd;
        const int* c = col4 + (size_t)i * 4;
        for (int s = 0; s < 4; ++s) {
            const bool used = c[s] >= 0;
            const float4 l = used ? nodeLo[c[s]] : make_float4(kFar, kFar, kFar, 0), h = used ? nodeHi[c[s]] : l;
            nd.box[0][s] = l.x; nd.box[1][s] = h.x; nd.box[2][s] = l.y;
            nd.box[3][s] = h.y; nd.box[4][s] = l.z; nd.box[5][s] = h.z;
            nd.child[s] = used ? childRef(c[s], n, leafMark, count, offset, idx4) : 0;
            nd.pad[s] = 0;
        }
        out4[idx4[j]] = nd;
    }
    if (st & 4) {
        Bvh8HostNode nd;
        const int* c = col8 + (size_t)i * 8;
        for (int s = 0; s < 8; ++s) {
            const bool used = c[s] >= 0;
            const float4 l = used ? nodeLo[c[s]] : make_float4(kFar, kFar, kFar, 0), h = used ? nodeHi[c[s]] : l;
            nd.box[0][s] = l.x; nd.box[1][s] = h.x; nd.box[2][s] = l.y;
            nd.box[3][s] = h.y; nd.box[4][s] = l.z; nd.box[5][s] = h.z;
            nd.child[s] = used ? childRef(c[s], n, leafMark, count, offset, idx8) : 0;
            nd.pad[s] = 0;
        }
        out8[idx8[j]] = nd;
    }
}

// Leaf-ordered triangles: 3 float4, v0.w = global triangle id bits.
__global__ void k_emit_tris(const BteTriangle* tris, const int* order, int n, float4* leafTris) {
    const int p = blockIdx.x * kB + threadIdx.x;
    if (p >= n) return;
    const int id = order[p];
    const BteTriangle t = tris[id];
    leafTris[3 * (size_t)p + 0] = make_float4(t.v0.x, t.v0.y, t.v0.z, __int_as_float(id));
    leafTris[3 * (size_t)p + 1] = make_float4(t.v1.x, t.v1.y, t.v1.z, 0.0f);
    leafTris[3 * (size_t)p + 2] = make_float4(t.v2.x, t.v2.y, t.v2.z, 0.0f);
}

// Every child reference of every node in range: inner nodes below `nodes`,
// leaves inside the triangle array.  Clears *ok otherwise.
__global__ void k_validate(const int* refs, int stride, int slots, int nodes, int ntris, int* ok) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= nodes) return;
    for (int s = 0; s < slots; ++s) {
        const int c = refs[(size_t)i * stride + s];
        if (c >= 0) {
            if (c >= nodes) atomicAnd(ok, 0);
        } else {
            const int code = ~c;
            const int start = code >> kLeafCountBits, cnt = (code & (kMaxLeafTris - 1)) + 1;
            if (start < 0 || start + cnt > ntris) atomicAnd(ok, 0);
        }
    }
}

__global__ void k_fill_i32(int* p, long long n, int v) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, long long n, uint8_t v) {
    const long long i = (long long)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}
// Root set-up: kept inner node, BVH4 and BVH8 node, depth 1, offset 0.
__global__ void k_root(int root, uint8_t* state, int* offset, int* depth2, int* depth4, int* depth8, int* maxDepths) {
    state[root] = 1 | 2 | 4;
    offset[root] = 0;
    depth2[root] = 1;
    depth4[root] = 1;
    depth8[root] = 1;
    maxDepths[0] = 1;
    maxDepths[1] = 1;
    maxDepths[2] = 1;
}

// Exclusive scan of n ints; total written to *total (device).
struct Scanner {
    DBuf tmp;
    int run(const int* in, int* out, int n, hipStream_t s) {
        size_t bytes = 0;
        HIPCHK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s));
        CHECK_RC(tmp.ensure(std::max<size_t>(bytes, 16)));
        HIPCHK(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s));
        return BTE_OK;
    }
};

}  // namespace

int build_bvh_gpu(const BteTriangle* dTris, int n, hipStream_t s, GpuBvh& out) {
    if (n < 2) return fail(BTE_ERR_UNSUPPORTED, "the GPU BVH builder needs at least 2 triangles");
    const auto t0 = std::chrono::steady_clock::now();
    const int total = 2 * n - 1, inner = n - 1;
    DBuf triLo, triHi, cen, bLo, bHi, keys, keys2, ids, ids2, nodeLo, nodeHi, ch0, ch1, clA, clB, nn, mf, mi, kf, ki;
    const int nb = blocks(n);
    CHECK_RC(triLo.ensure((size_t)n * 16));
    CHECK_RC(triHi.ensure((size_t)n * 16));
    CHECK_RC(cen.ensure((size_t)n * 16));
    CHECK_RC(bLo.ensure((size_t)std::max(nb, kB + 1) * 16));
    CHECK_RC(bHi.ensure((size_t)std::max(nb, 1) * 16));
    hipLaunchKernelGGL(k_tri_boxes, dim3(nb), dim3(kB), 0, s, dTris, n, triLo.as<float4>(), triHi.as<float4>(),
                       cen.as<float4>(), bLo.as<float4>(), bHi.as<float4>());
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kB), 0, s, bLo.as<float4>(), bHi.as<float4>(), nb);
    // the high corner goes next to the low one (bounds[kB]) for k_morton
    HIPCHK(hipMemcpyAsync(bLo.as<float4>() + kB, bHi.p, 16, hipMemcpyDeviceToDevice, s));
    CHECK_RC(keys.ensure((size_t)n * 8));
    CHECK_RC(keys2.ensure((size_t)n * 8));
    CHECK_RC(ids.ensure((size_t)n * 4));
    CHECK_RC(ids2.ensure((size_t)n * 4));
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kB), 0, s, cen.as<float4>(), n, bLo.as<float4>(), keys.as<uint64_t>(),
                       ids.as<int>());
    {
        size_t bytes = 0;
        HIPCHK(rocprim::radix_sort_pairs(nullptr, bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), ids.as<int>(),
                                         ids2.as<int>(), n, 0, 63, s));
        DBuf tmp;
        CHECK_RC(tmp.ensure(std::max<size_t>(bytes, 16)));
        HIPCHK(rocprim::radix_sort_pairs(tmp.p, bytes, keys.as<uint64_t>(), keys2.as<uint64_t>(), ids.as<int>(),
                                         ids2.as<int>(), n, 0, 63, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    keys.release();
    keys2.release();
    cen.release();
    const int* sortedIds = ids2.as<int>();
    CHECK_RC(nodeLo.ensure((size_t)total * 16));
    CHECK_RC(nodeHi.ensure((size_t)total * 16));
    CHECK_RC(ch0.ensure((size_t)total * 4));
    CHECK_RC(ch1.ensure((size_t)total * 4));
    CHECK_RC(clA.ensure((size_t)n * 4));
    CHECK_RC(clB.ensure((size_t)n * 4));
    for (DBuf* b : {&nn, &mf, &mi, &kf, &ki}) CHECK_RC(b->ensure((size_t)n * 4));
    hipLaunchKernelGGL(k_leaf_nodes, dim3(nb), dim3(kB), 0, s, sortedIds, n, triLo.as<float4>(), triHi.as<float4>(),
                       nodeLo.as<float4>(), nodeHi.as<float4>(), clA.as<int>());
    triLo.release();
    triHi.release();

    // clustering rounds
    Scanner scan;
    std::vector<int> roundFirst, roundCount;
    int m = n, nodeNext = n;
    int* cur = clA.as<int>();
    int* nxt = clB.as<int>();
    const int kMaxRounds = 1 << 16;
    while (m > 1) {
        if ((int)roundFirst.size() >= kMaxRounds) return fail(BTE_ERR_UNSUPPORTED, "BVH clustering did not converge");
        const int mb = blocks(m);
        hipLaunchKernelGGL(k_nearest, dim3(mb), dim3(kB), 0, s, cur, m, nodeLo.as<float4>(), nodeHi.as<float4>(),
                           nn.as<int>());
        hipLaunchKernelGGL(k_merge_flags, dim3(mb), dim3(kB), 0, s, nn.as<int>(), m, mf.as<int>(), kf.as<int>());
        CHECK_RC(scan.run(mf.as<int>(), mi.as<int>(), m, s));
        CHECK_RC(scan.run(kf.as<int>(), ki.as<int>(), m, s));
        int last[4];
        HIPCHK(hipMemcpyAsync(&last[0], mi.as<int>() + m - 1, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&last[1], mf.as<int>() + m - 1, 4, hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_merge_apply, dim3(mb), dim3(kB), 0, s, cur, nn.as<int>(), m, mf.as<int>(), mi.as<int>(),
                           kf.as<int>(), ki.as<int>(), nodeNext, ch0.as<int>(), ch1.as<int>(), nodeLo.as<float4>(),
                           nodeHi.as<float4>(), nxt);
        HIPCHK(hipStreamSynchronize(s));
        const int merges = last[0] + last[1];
        if (merges <= 0) return fail(BTE_ERR_DEVICE, "BVH clustering round without a merge");
        roundFirst.push_back(nodeNext);
        roundCount.push_back(merges);
        nodeNext += merges;
        m -= merges;
        std::swap(cur, nxt);
    }
    const int root = nodeNext - 1;   // == 2n - 2
    for (DBuf* b : {&clA, &clB, &nn, &mf, &mi, &kf, &ki}) b->release();

    // bottom-up: counts, BVH2 leaves, wide-collapse dynamic programs
    DBuf count, cost2, leafMark, cost4, split4, cost8, split8;
    CHECK_RC(count.ensure((size_t)total * 4));
    CHECK_RC(cost2.ensure((size_t)total * 4));
    CHECK_RC(leafMark.ensure((size_t)total));
    CHECK_RC(cost4.ensure((size_t)total * 5 * 4));
    CHECK_RC(split4.ensure((size_t)total * 5));
    CHECK_RC(cost8.ensure((size_t)total * 9 * 4));
    CHECK_RC(split8.ensure((size_t)total * 9));
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks(total)), dim3(kB), 0, s, leafMark.as<uint8_t>(), (long long)total,
                       (uint8_t)0);
    for (size_t r = 0; r < roundFirst.size(); ++r)
        hipLaunchKernelGGL(k_bottom_up, dim3(blocks(roundCount[r])), dim3(kB), 0, s, roundFirst[r], roundCount[r], n,
                           root, ch0.as<int>(), ch1.as<int>(), nodeLo.as<float4>(), nodeHi.as<float4>(),
                           count.as<int>(), cost2.as<float>(), leafMark.as<uint8_t>(), cost4.as<float>(),
                           split4.as<int8_t>(), cost8.as<float>(), split8.as<int8_t>());
    cost2.release();

    // top-down: leaf ranges, depths, wide children
    DBuf offset, depth2, depth4, depth8, state, order, col4, col8, maxDepths;
    CHECK_RC(offset.ensure((size_t)total * 4));
    CHECK_RC(depth2.ensure((size_t)total * 4));
    CHECK_RC(depth4.ensure((size_t)total * 4));
    CHECK_RC(depth8.ensure((size_t)total * 4));
    CHECK_RC(state.ensure((size_t)total));
    CHECK_RC(order.ensure((size_t)n * 4));
    CHECK_RC(col4.ensure((size_t)total * 4 * 4));
    CHECK_RC(col8.ensure((size_t)total * 8 * 4));
    CHECK_RC(maxDepths.ensure(16));   // 3 depths + the validation flag
    hipLaunchKernelGGL(k_fill_u8, dim3(blocks(total)), dim3(kB), 0, s, state.as<uint8_t>(), (long long)total, (uint8_t)0);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(total)), dim3(kB), 0, s, offset.as<int>(), (long long)total, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(kB), 0, s, order.as<int>(), (long long)n, 0);
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, s, root, state.as<uint8_t>(), offset.as<int>(), depth2.as<int>(),
                       depth4.as<int>(), depth8.as<int>(), maxDepths.as<int>());
    for (size_t r = roundFirst.size(); r-- > 0;)
        hipLaunchKernelGGL(k_top_down, dim3(blocks(roundCount[r])), dim3(kB), 0, s, roundFirst[r], roundCount[r], n,
                           ch0.as<int>(), ch1.as<int>(), leafMark.as<uint8_t>(), count.as<int>(), split4.as<int8_t>(),
                           split8.as<int8_t>(), offset.as<int>(), depth2.as<int>(), depth4.as<int>(), depth8.as<int>(),
                           state.as<uint8_t>(), sortedIds, order.as<int>(), col4.as<int>(), col8.as<int>(),
                           maxDepths.as<int>());
    for (DBuf* b : {&cost4, &split4, &cost8, &split8, &depth2, &depth4, &depth8}) b->release();

    // numbering (root first) and emission
    DBuf f2, f4, f8, i2, i4, i8;
    for (DBuf* b : {&f2, &f4, &f8, &i2, &i4, &i8}) CHECK_RC(b->ensure((size_t)inner * 4));
    hipLaunchKernelGGL(k_number_flags, dim3(blocks(inner)), dim3(kB), 0, s, n, inner, state.as<uint8_t>(), f2.as<int>(),
                       f4.as<int>(), f8.as<int>());
    CHECK_RC(scan.run(f2.as<int>(), i2.as<int>(), inner, s));
    CHECK_RC(scan.run(f4.as<int>(), i4.as<int>(), inner, s));
    CHECK_RC(scan.run(f8.as<int>(), i8.as<int>(), inner, s));
    int tail[6], depths[3];
    HIPCHK(hipMemcpyAsync(&tail[0], i2.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&tail[1], f2.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&tail[2], i4.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&tail[3], f4.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&tail[4], i8.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&tail[5], f8.as<int>() + inner - 1, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(depths, maxDepths.p, 12, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    out.nodes2 = tail[0] + tail[1];
    out.nodes4 = tail[2] + tail[3];
    out.nodes8 = tail[4] + tail[5];
    out.depth2 = depths[0];
    out.depth4 = depths[1];
    out.depth8 = depths[2];
    out.rounds = (int)roundFirst.size();
    CHECK_RC(out.dNodes2.ensure((size_t)out.nodes2 * sizeof(BvhHostNode)));
    CHECK_RC(out.dNodes4.ensure((size_t)out.nodes4 * sizeof(Bvh4HostNode)));
    CHECK_RC(out.dNodes8.ensure((size_t)out.nodes8 * sizeof(Bvh8HostNode)));
    CHECK_RC(out.dLeafTris.ensure((size_t)n * 3 * sizeof(float4)));
    hipLaunchKernelGGL(k_emit_nodes, dim3(blocks(inner)), dim3(kB), 0, s, n, inner, ch0.as<int>(), ch1.as<int>(),
                       leafMark.as<uint8_t>(), count.as<int>(), offset.as<int>(), state.as<uint8_t>(),
                       nodeLo.as<float4>(), nodeHi.as<float4>(), col4.as<int>(), col8.as<int>(), i2.as<int>(),
                       i4.as<int>(), i8.as<int>(), out.dNodes2.as<BvhHostNode>(), out.dNodes4.as<Bvh4HostNode>(),
                       out.dNodes8.as<Bvh8HostNode>());
    hipLaunchKernelGGL(k_emit_tris, dim3(nb), dim3(kB), 0, s, dTris, order.as<int>(), n, out.dLeafTris.as<float4>());
    // a reference outside the arrays would make a traversal read out of
    // bounds: check every one before the trees are used
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kB), 0, s, maxDepths.as<int>() + 3, 1ll, 1);
    hipLaunchKernelGGL(k_validate, dim3(blocks(out.nodes2)), dim3(kB), 0, s,
                       reinterpret_cast<const int*>(out.dNodes2.p) + 12, 16, 2, out.nodes2, n, maxDepths.as<int>() + 3);
    hipLaunchKernelGGL(k_validate, dim3(blocks(out.nodes4)), dim3(kB), 0, s,
                       reinterpret_cast<const int*>(out.dNodes4.p) + 24, 32, 4, out.nodes4, n, maxDepths.as<int>() + 3);
    hipLaunchKernelGGL(k_validate, dim3(blocks(out.nodes8)), dim3(kB), 0, s,
                       reinterpret_cast<const int*>(out.dNodes8.p) + 48, 64, 8, out.nodes8, n, maxDepths.as<int>() + 3);
    int valid = 0;
    HIPCHK(hipMemcpyAsync(&valid, maxDepths.as<int>() + 3, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    if (!valid) return fail(BTE_ERR_DEVICE, "GPU BVH build produced an out-of-range child reference");
    out.buildMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BTE_OK;
}

}  // namespace bte
