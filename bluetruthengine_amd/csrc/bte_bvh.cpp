// bte_bvh.cpp — BVH2 builder (replaces optixAccelBuild, src/host/optix_setup.cu:215-292).
//
// Binned SAH (16 bins per axis over centroid bounds), leaves of at most
// kMaxLeafTris triangles, object-median splits once the depth budget requires
// them so the tree never exceeds kMaxBvhDepth levels (the traversal stack
// holds kStackSize entries).  Leaves reference a leaf-ordered copy of the
// triangles whose v0.w carries the global triangle index, so hit IDs are the
// flattened scene indices (the invariant of docs/MEGAKERNEL_WALKTHROUGH.md:158-163).
#include "bte_bvh.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bte {

namespace {

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    }
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void growPoint(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    std::vector<Box> triBox;
    std::vector<float> cent;   // 3 per triangle
    std::vector<int> order;
    std::vector<BvhHostNode> nodes;
    int maxDepth = 0;

    static int ceilLog2(int n) {
        int l = 0;
        while ((1 << l) < n) ++l;
        return l;
    }

    Box rangeBox(int b, int e) const {
        Box bx;
        bx.reset();
        for (int i = b; i < e; ++i) bx.grow(triBox[order[i]]);
        return bx;
    }

    static int leafCode(int start, int count) { return ~((start << kLeafCountBits) | (count - 1)); }

    // Returns the reference (node index or leaf code) of the subtree over [b, e).
    int build(int b, int e, int depth, Box& outBox) {
        const int n = e - b;
        outBox = rangeBox(b, e);
        maxDepth = std::max(maxDepth, depth);
        if (n <= 2 && n <= kMaxLeafTris) return leafCode(b, n);

        Box cb;
        cb.reset();
        for (int i = b; i < e; ++i) cb.growPoint(&cent[3 * (size_t)order[i]]);
        int axis = 0;
        float ext = cb.hi[0] - cb.lo[0];
        for (int k = 1; k < 3; ++k)
            if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; axis = k; }

        int mid = -1;
        const bool forceMedian = depth + ceilLog2(n) + 1 >= kMaxBvhDepth;
        if (!forceMedian && ext > 0.0f) {
            constexpr int NB = 16;
            float bestCost = INFINITY;
            int bestAxis = -1, bestSplit = -1;
            for (int a = 0; a < 3; ++a) {
                const float lo = cb.lo[a], hi = cb.hi[a];
                if (!(hi > lo)) continue;
                const float scale = NB / (hi - lo);
                Box bins[NB];
                int cnt[NB] = {0};
                for (int k = 0; k < NB; ++k) bins[k].reset();
                for (int i = b; i < e; ++i) {
                    const int t = order[i];
                    int bi = (int)((cent[3 * (size_t)t + a] - lo) * scale);
                    bi = std::min(std::max(bi, 0), NB - 1);
                    bins[bi].grow(triBox[t]);
                    cnt[bi]++;
                }
                float rightArea[NB];
                int rightCnt[NB];
                Box acc;
                acc.reset();
                int c = 0;
                for (int k = NB - 1; k > 0; --k) {
                    acc.grow(bins[k]);
                    c += cnt[k];
                    rightArea[k] = acc.area();
                    rightCnt[k] = c;
                }
                acc.reset();
                c = 0;
                for (int k = 0; k < NB - 1; ++k) {
                    acc.grow(bins[k]);
                    c += cnt[k];
                    if (c == 0 || rightCnt[k + 1] == 0) continue;
                    const float cost = acc.area() * c + rightArea[k + 1] * rightCnt[k + 1];
                    if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = k; }
                }
            }
            if (bestAxis >= 0) {
                const float parentArea = outBox.area();
                const float splitCost = 1.0f + (parentArea > 0.0f ? bestCost / parentArea : 0.0f);
                if (n <= kMaxLeafTris && splitCost >= (float)n) return leafCode(b, n);
                const float lo = cb.lo[bestAxis];
                const float scale = NB / (cb.hi[bestAxis] - lo);
                int* it = std::partition(order.data() + b, order.data() + e, [&](int t) {
                    int bi = (int)((cent[3 * (size_t)t + bestAxis] - lo) * scale);
                    bi = std::min(std::max(bi, 0), NB - 1);
                    return bi <= bestSplit;
                });
                mid = (int)(it - order.data());
                if (mid == b || mid == e) mid = -1;
            }
        }
        if (mid < 0) {
            if (n <= kMaxLeafTris) return leafCode(b, n);
            mid = b + n / 2;
            std::nth_element(order.data() + b, order.data() + mid, order.data() + e, [&](int x, int y) {
                const float cx = cent[3 * (size_t)x + axis], cy = cent[3 * (size_t)y + axis];
                return cx < cy || (cx == cy && x < y);
            });
        }
        const int idx = (int)nodes.size();
        nodes.push_back(BvhHostNode{});
        Box lb, rb;
        const int l = build(b, mid, depth + 1, lb);
        const int r = build(mid, e, depth + 1, rb);
        BvhHostNode& nd = nodes[idx];
        nd.n[0] = lb.lo[0]; nd.n[1] = lb.hi[0]; nd.n[2] = lb.lo[1]; nd.n[3] = lb.hi[1];
        nd.n[4] = rb.lo[0]; nd.n[5] = rb.hi[0]; nd.n[6] = rb.lo[1]; nd.n[7] = rb.hi[1];
        nd.n[8] = lb.lo[2]; nd.n[9] = lb.hi[2]; nd.n[10] = rb.lo[2]; nd.n[11] = rb.hi[2];
        nd.child[0] = l;
        nd.child[1] = r;
        nd.pad[0] = nd.pad[1] = 0;
        return idx;
    }
};

}  // namespace

BvhBuild build_bvh2(const BteTriangle* tris, int n) {
    BvhBuild out;
    Builder bl;
    bl.triBox.resize((size_t)n);
    bl.cent.resize(3 * (size_t)n);
    bl.order.resize((size_t)n);
    for (int i = 0; i < n; ++i) {
        const BteTriangle& t = tris[i];
        Box b;
        b.reset();
        const float v[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
        for (int j = 0; j < 3; ++j) b.growPoint(v[j]);
        bl.triBox[i] = b;
        for (int k = 0; k < 3; ++k) bl.cent[3 * (size_t)i + k] = (b.lo[k] + b.hi[k]) * 0.5f;
        bl.order[i] = i;
    }
    bl.nodes.reserve(n > 0 ? (size_t)(2 * n / 2 + 1) : 1);
    if (n == 0) {
        BvhHostNode root{};
        for (int k = 0; k < 12; ++k) root.n[k] = -3.0e38f;
        root.child[0] = root.child[1] = ~0;
        bl.nodes.push_back(root);
    } else {
        Box rootBox;
        bl.nodes.push_back(BvhHostNode{});   // reserve the root slot
        // Build the tree under a temporary root, then hoist: node 0 must be inner.
        const int ref = bl.build(0, n, 1, rootBox);
        if (ref >= 0) {
            bl.nodes[0] = bl.nodes[ref];
            // node `ref` stays as an unreferenced duplicate; harmless
        } else {
            BvhHostNode& r = bl.nodes[0];
            r.n[0] = rootBox.lo[0]; r.n[1] = rootBox.hi[0]; r.n[2] = rootBox.lo[1]; r.n[3] = rootBox.hi[1];
            r.n[8] = rootBox.lo[2]; r.n[9] = rootBox.hi[2];
            r.n[4] = r.n[5] = r.n[6] = r.n[7] = r.n[10] = r.n[11] = -3.0e38f;   // empty, far away
            r.child[0] = ref;
            r.child[1] = ref;
            r.pad[0] = r.pad[1] = 0;
        }
    }
    out.nodes = std::move(bl.nodes);
    out.depth = bl.maxDepth;
    out.leafTris.resize(12 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const int id = bl.order[i];
        const BteTriangle& t = tris[id];
        float* p = &out.leafTris[12 * (size_t)i];
        p[0] = t.v0.x; p[1] = t.v0.y; p[2] = t.v0.z; std::memcpy(&p[3], &id, 4);
        p[4] = t.v1.x; p[5] = t.v1.y; p[6] = t.v1.z; p[7] = 0.0f;
        p[8] = t.v2.x; p[9] = t.v2.y; p[10] = t.v2.z; p[11] = 0.0f;
    }
    return out;
}

namespace {

struct Child2 {
    int ref;
    float lo[3], hi[3];
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return (dx < 0.0f || dy < 0.0f || dz < 0.0f) ? -1.0f : 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

Child2 childOf(const BvhHostNode& n, int k) {
    Child2 c;
    c.ref = n.child[k];
    if (k == 0) {
        c.lo[0] = n.n[0]; c.hi[0] = n.n[1]; c.lo[1] = n.n[2]; c.hi[1] = n.n[3]; c.lo[2] = n.n[8]; c.hi[2] = n.n[9];
    } else {
        c.lo[0] = n.n[4]; c.hi[0] = n.n[5]; c.lo[1] = n.n[6]; c.hi[1] = n.n[7]; c.lo[2] = n.n[10]; c.hi[2] = n.n[11];
    }
    return c;
}

// Collapse by surface-area cost (dynamic programming over the BVH2, after
// Ylitie, Karras and Laine, "Efficient incoherent ray traversal on GPUs
// through compressed wide BVHs", HPG 2017): for every BVH2 inner node and a
// budget of 1..W child slots, the cheapest way to represent its subtree,
// either as one W-wide node (one slot) or spread over the slots through its
// two children.  A node visit costs kNodeCost per unit of box area, a leaf
// kTriCost per triangle per unit area; BVH2 leaves are kept as they are.
template <int W, class Node>
struct SahCollapser {
    static constexpr float kNodeCost = 1.0f;
    static constexpr float kTriCost = 0.3f;
    static constexpr int kS = W + 1;
    const std::vector<BvhHostNode>& n2;
    std::vector<float> cost;     // [node * kS + slots], slots 1..W
    std::vector<int8_t> split;   // [node * kS + slots]: slots given to child 0 (0 = one node), slot 0: node's own split
    std::vector<Node> out;
    int maxDepth = 0;

    explicit SahCollapser(const std::vector<BvhHostNode>& nodes) : n2(nodes) {}

    float childCost(const Child2& c, int slots) const {
        if (c.ref < 0) {
            const int cnt = ((~c.ref) & (kMaxLeafTris - 1)) + 1;
            return std::max(c.area(), 0.0f) * kTriCost * (float)cnt;
        }
        return cost[(size_t)c.ref * kS + (size_t)slots];
    }
    float distribute(int i, int slots, int& bestK) const {
        const Child2 a = childOf(n2[(size_t)i], 0), b = childOf(n2[(size_t)i], 1);
        float best = __builtin_huge_valf();
        bestK = 1;
        for (int k = 1; k < slots; ++k) {
            const float c = childCost(a, k) + childCost(b, slots - k);
            if (c < best) {
                best = c;
                bestK = k;
            }
        }
        return best;
    }
    void solve() {
        const size_t n = n2.size();
        cost.assign(n * kS, 0.0f);
        split.assign(n * kS, 0);
        // parents precede their children in the BVH2 array
        for (size_t i = n; i-- > 0;) {
            const BvhHostNode& nd = n2[i];
            if (nd.child[0] == nd.child[1]) {   // single-leaf root: its second box is the far point
                for (int k = 1; k <= W; ++k) cost[i * kS + (size_t)k] = 0.0f;
                continue;
            }
            const Child2 a = childOf(nd, 0), b = childOf(nd, 1);
            Child2 u = a;
            for (int ax = 0; ax < 3; ++ax) {
                u.lo[ax] = std::min(a.lo[ax], b.lo[ax]);
                u.hi[ax] = std::max(a.hi[ax], b.hi[ax]);
            }
            int kw = 1;
            const float asNode = std::max(u.area(), 0.0f) * kNodeCost + distribute((int)i, W, kw);
            cost[i * kS + 1] = asNode;
            split[i * kS + 0] = (int8_t)kw;
            for (int slots = 2; slots <= W; ++slots) {
                int k = 1;
                const float spread = distribute((int)i, slots, k);
                cost[i * kS + (size_t)slots] = std::min(asNode, spread);
                split[i * kS + (size_t)slots] = spread < asNode ? (int8_t)k : (int8_t)0;
            }
        }
    }
    void collect(const Child2& c, int slots, Child2* acc, int& cnt) const {
        if (c.ref < 0 || slots == 1 || split[(size_t)c.ref * kS + (size_t)slots] == 0) {
            acc[cnt++] = c;
            return;
        }
        const int k = split[(size_t)c.ref * kS + (size_t)slots];
        const BvhHostNode& nd = n2[(size_t)c.ref];
        collect(childOf(nd, 0), k, acc, cnt);
        collect(childOf(nd, 1), slots - k, acc, cnt);
    }
    // W-wide node for BVH2 inner node `i`; returns its index
    int make(int i, int depth) {
        maxDepth = std::max(maxDepth, depth);
        Child2 ch[W];
        int cnt = 0;
        const BvhHostNode& root = n2[(size_t)i];
        if (root.child[0] == root.child[1]) {
            ch[cnt++] = childOf(root, 0);
        } else {
            const int k = split[(size_t)i * kS + 0];
            collect(childOf(root, 0), k, ch, cnt);
            collect(childOf(root, 1), W - k, ch, cnt);
        }
        const int idx = (int)out.size();
        out.push_back(Node{});
        int refs[W] = {};
        for (int k = 0; k < cnt; ++k) refs[k] = ch[k].ref >= 0 ? make(ch[k].ref, depth + 1) : ch[k].ref;
        Node& nd = out[(size_t)idx];
        for (int k = 0; k < W; ++k) {
            const bool used = k < cnt;
            for (int a = 0; a < 3; ++a) {
                nd.box[2 * a][k] = used ? ch[k].lo[a] : -3.0e38f;
                nd.box[2 * a + 1][k] = used ? ch[k].hi[a] : -3.0e38f;
            }
            nd.child[k] = refs[k];
            nd.pad[k] = 0;
        }
        return idx;
    }
};

}  // namespace

std::vector<Bvh4HostNode> collapse_bvh4(const std::vector<BvhHostNode>& nodes2, int& depth4) {
    SahCollapser<4, Bvh4HostNode> c(nodes2);
    c.solve();
    c.out.reserve(nodes2.size() / 2 + 1);
    c.make(0, 1);
    depth4 = c.maxDepth;
    return std::move(c.out);
}

std::vector<Bvh8HostNode> collapse_bvh8(const std::vector<BvhHostNode>& nodes2, int& depth8) {
    SahCollapser<8, Bvh8HostNode> c(nodes2);
    c.solve();
    c.out.reserve(nodes2.size() / 4 + 1);
    c.make(0, 1);
    depth8 = c.maxDepth;
    return std::move(c.out);
}

}  // namespace bte
