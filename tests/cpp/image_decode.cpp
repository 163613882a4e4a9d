// Decodes a PNG or JPEG file with the library's host decoders (bte_image.cpp,
// bte_png.cpp) and writes the width, the height (int32) and the RGBA8 pixels
// (row 0 = top) to a raw file:
//     image_decode in.png|in.jpg out.rgba
#include <cstdio>
#include <vector>

#include "bte_image.h"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<uint8_t> data;
    int c;
    while ((c = std::fgetc(f)) != EOF) data.push_back((uint8_t)c);
    std::fclose(f);
    int w = 0, h = 0;
    std::vector<uint8_t> px;
    std::string err;
    if (!bte::decode_image(data.data(), data.size(), w, h, px, err)) {
        std::printf("error: %s\n", err.c_str());
        return 1;
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    const int dims[2] = {w, h};
    std::fwrite(dims, sizeof(int), 2, o);
    std::fwrite(px.data(), 1, px.size(), o);
    std::fclose(o);
    std::printf("%d %d\n", w, h);
    return 0;
}
