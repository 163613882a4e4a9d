// Headless counterpart of the reference's main(): loads a scene.xml through
// the C++ host layer, renders `frames` wavefront frames into the library's
// display buffer and writes it as a binary PPM (top row first).
// Usage: scene_render scene.xml frames out.ppm
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bte_render_kernel.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s scene.xml frames out.ppm\n", argv[0]);
        return 2;
    }
    const BteWindowSize win = loadSceneXml(argv[1]);
    cudaRegisterPBO(nullptr);
    cudaSetRenderMode(RenderMode::Wavefront);
    cudaResetAccumulation(win.width, win.height);
    const int frames = std::atoi(argv[2]);
    for (int f = 0; f < frames; ++f) cudaRender(win.width, win.height);
    std::vector<uint32_t> px((size_t)win.width * win.height);
    bte_compat::check(bte_read_framebuffer(px.data(), px.size()), "bte_read_framebuffer");
    FILE* out = std::fopen(argv[3], "wb");
    if (!out) return 3;
    std::fprintf(out, "P6\n%d %d\n255\n", win.width, win.height);
    for (int y = win.height - 1; y >= 0; --y)   // row 0 of the buffer is the bottom row
        for (int x = 0; x < win.width; ++x) {
            const uint32_t p = px[(size_t)y * win.width + x];
            const unsigned char rgb[3] = {(unsigned char)(p & 255), (unsigned char)((p >> 8) & 255),
                                          (unsigned char)((p >> 16) & 255)};
            std::fwrite(rgb, 1, 3, out);
        }
    std::fclose(out);
    std::printf("rendered %d frames of %dx%d\n", frames, win.width, win.height);
    cudaCleanup();
    return 0;
}
