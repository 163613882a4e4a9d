// Builds a one-triangle scene through the reference-named C++ host layer
// (include/bte_render_kernel.hpp) and renders one frame; prints the number of
// lit pixels.  Compiled by tests/test_cpp_compat.py.
#include <cstdio>
#include <vector>

#include "bte_render_kernel.hpp"

int main() {
    const int W = 32, H = 32;
    TriangleData tri{};
    tri.v0 = {-1.0f, -1.0f, 0.0f};
    tri.v1 = {1.0f, -1.0f, 0.0f};
    tri.v2 = {0.0f, 1.0f, 0.0f};
    tri.n0 = tri.n1 = tri.n2 = {0.0f, 0.0f, 1.0f};
    cudaInit(tri, Float3{1.0f, 1.0f, 1.0f}, W, H);
    CameraData cam{};
    cam.origin = {0.0f, 0.0f, 3.0f};
    cam.forward = {0.0f, 0.0f, -1.0f};
    cam.right = {1.0f, 0.0f, 0.0f};
    cam.up = {0.0f, 1.0f, 0.0f};
    cam.fovYRadians = 0.9f;
    cam.aspect = 1.0f;
    cudaInitCamera(cam);
    const Float3 emission{2.0f, 2.0f, 2.0f};
    const uint8_t flag = 1;
    cudaInitTriangleEmission(&emission, 1);
    cudaInitTriangleEmitterFlags(&flag, 1);
    BsdfData b{};
    b.type = BSDF_Diffuse;
    b.p0 = {0.5f, 0.5f, 0.5f, 0.0f};
    const int bid = 0;
    cudaInitBsdfs(&b, 1, &bid, 1);
    cudaInitOptix();
    cudaRegisterPBO(nullptr);
    cudaSetRenderMode(RenderMode::Wavefront);
    cudaRender(W, H);
    std::vector<float> acc(3 * W * H);
    if (bte_read_accum(acc.data(), acc.size()) != BTE_OK) return 2;
    int lit = 0;
    for (int i = 0; i < W * H; ++i) lit += acc[3 * i] == 2.0f;
    std::printf("lit=%d\n", lit);
    cudaCleanup();
    return lit > 0 ? 0 : 1;
}
