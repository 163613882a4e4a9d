// Structural check of the BVH4 and BVH8 built by the library's host code
// (BVH2 SAH build, then the surface-area-cost collapse): over random triangle
// sets (flat ones included) every triangle is referenced by exactly one leaf,
// each child box contains everything below it, at most W children per node,
// and the collapsed depth stays within the BVH2 depth (the traversal stack is
// sized from it).  Prints one line per set and exits 0.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "bte_bvh.h"

using namespace bte;

// Walks a W-wide BVH from the root; returns an error message or "".
template <int W, typename Node>
static std::string checkWide(const std::vector<Node>& nw, int depthW, const BvhBuild& b, int n) {
    std::vector<int> seen((size_t)n, 0);
    long long nodes = 0;
    // returns false on a containment failure; lo/hi: the box the parent gave this node
    std::function<bool(int, const float*, const float*)> walk = [&](int node, const float* plo, const float* phi) {
        ++nodes;
        const Node& nd = nw[(size_t)node];
        for (int k = 0; k < W; ++k) {
            if (nd.box[0][k] == -3.0e38f && nd.box[1][k] == -3.0e38f) continue;
            const float lo[3] = {nd.box[0][k], nd.box[2][k], nd.box[4][k]};
            const float hi[3] = {nd.box[1][k], nd.box[3][k], nd.box[5][k]};
            if (plo)
                for (int a = 0; a < 3; ++a)
                    if (lo[a] < plo[a] || hi[a] > phi[a]) return false;
            const int c = nd.child[k];
            if (c >= 0) {
                if (!walk(c, lo, hi)) return false;
                continue;
            }
            const int code = ~c, start = code >> kLeafCountBits, cnt = (code & (kMaxLeafTris - 1)) + 1;
            for (int i = start; i < start + cnt; ++i) {
                const float* p = &b.leafTris[12 * (size_t)i];
                int id;
                std::memcpy(&id, &p[3], 4);
                seen[(size_t)id]++;
                for (int v = 0; v < 3; ++v)
                    for (int a = 0; a < 3; ++a)
                        if (p[4 * v + a] < lo[a] || p[4 * v + a] > hi[a]) return false;
            }
        }
        return true;
    };
    if (!walk(0, nullptr, nullptr)) return "a box does not contain its subtree";
    for (int i = 0; i < n; ++i)
        if (seen[(size_t)i] != 1) return "triangle " + std::to_string(i) + " referenced " + std::to_string(seen[(size_t)i]) + " times";
    if (nodes != (long long)nw.size() || depthW > b.depth || depthW > kMaxBvhDepth)
        return std::to_string(nodes) + " reachable of " + std::to_string(nw.size()) + " nodes, depth " +
               std::to_string(depthW) + " (BVH2 " + std::to_string(b.depth) + ")";
    return "";
}

int main() {
    std::mt19937 rng(3);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    for (int trial = 0; trial < 6; ++trial) {
        const int n = trial == 0 ? 1 : (trial == 1 ? 7 : 5000 + trial * 7777);
        std::vector<BteTriangle> tris((size_t)n);
        for (auto& t : tris) {
            std::memset(&t, 0, sizeof t);
            float* v = reinterpret_cast<float*>(&t);
            const float c[3] = {U(rng) * 4, U(rng) * (trial & 1 ? 0.01f : 2.0f), U(rng) * 4};
            for (int k = 0; k < 9; ++k) v[k] = c[k % 3] + 0.05f * U(rng);
        }
        BvhBuild b = build_bvh2(tris.data(), n);
        int depth4 = 0;
        const std::vector<Bvh4HostNode> n4 = collapse_bvh4(b.nodes, depth4);
        int depth8 = 0;
        const std::vector<Bvh8HostNode> n8 = collapse_bvh8(b.nodes, depth8);
        std::string e4 = checkWide<4>(n4, depth4, b, n), e8 = checkWide<8>(n8, depth8, b, n);
        if (!e4.empty() || !e8.empty()) {
            std::printf("set %d: BVH4: %s; BVH8: %s\n", trial, e4.c_str(), e8.c_str());
            return 1;
        }
        std::printf("set %d: %d triangles, BVH2 %zu nodes depth %d, BVH4 %zu nodes depth %d, BVH8 %zu nodes depth %d\n",
                    trial, n, b.nodes.size(), b.depth, n4.size(), depth4, n8.size(), depth8);
    }
    return 0;
}
