// bte_image.h — image decoding for the scene host layer (texture maps):
// baseline JPEG in bte_image.cpp, zlib / PNG in bte_png.cpp.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bte {

// Baseline sequential JPEG (ITU-T T.81: 8-bit samples, Huffman coding,
// interleaved or single-component scans, any sampling factors, restart
// intervals; grayscale or YCbCr) to RGBA8, row 0 = top.  Returns false with a
// reason for anything else (progressive, arithmetic, 12-bit, CMYK).
bool decode_jpeg(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                 std::string& err);

// zlib stream (RFC 1950 / 1951) -> bytes; checksums are not verified.
bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>& out, std::string& err);

// PNG (grey, RGB, palette, grey + alpha, RGBA; 1-16 bits; Adam7) to RGBA8,
// row 0 = top, with stb_image's 8-bit 4-channel conversions (high byte of
// 16-bit samples, scaled low-depth grey, tRNS colour key and palette alpha).
bool decode_png(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                std::string& err);

// PNG or baseline JPEG by signature.
bool decode_image(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                  std::string& err);

}  // namespace bte
