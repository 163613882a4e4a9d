// bte_image.h — image decoding for the scene host layer (texture maps).
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

}  // namespace bte
