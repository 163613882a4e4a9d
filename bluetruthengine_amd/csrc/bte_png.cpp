// bte_png.cpp — PNG decoding for texture maps (glTF images and scene.xml
// albedo textures).
//
// The reference decodes images with stb_image (src/main.cpp:442-475 for
// scene.xml textures; tinygltf's image callback for glTF images, which asks
// for 4 channels).  This file implements the formats from their
// specifications: zlib (RFC 1950) around DEFLATE (RFC 1951: stored, fixed and
// dynamic Huffman blocks), and PNG (ISO/IEC 15948: IHDR / PLTE / tRNS / IDAT
// chunks, the five scanline filters, Adam7 interlacing, bit depths 1-16, grey,
// RGB, palette, grey + alpha and RGBA).  Conversion to RGBA8 follows what an
// 8-bit, 4-channel stb_image load returns: 16-bit samples keep their high
// byte, grey samples below 8 bits are scaled to 0..255, a tRNS colour key
// gives alpha 0 to matching pixels, palette entries take their tRNS alpha.
// Checksums (CRC-32, Adler-32) are not verified, as stb_image does not.
#include "bte_image.h"

#include <cstring>

namespace bte {
namespace {

struct BitReader {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t acc = 0;
    int bits = 0;
    bool overrun = false;
    BitReader(const uint8_t* d, size_t s) : p(d), n(s) {}
    void fill() {
        while (bits <= 24) {
            uint32_t b = 0;
            if (pos < n) b = p[pos];
            else overrun = pos > n + 4;   // a few zero bytes past the end are allowed for peeking
            ++pos;
            acc |= b << bits;
            bits += 8;
        }
    }
    uint32_t peek(int k) {
        if (bits < k) fill();
        return acc & ((1u << k) - 1u);
    }
    void drop(int k) {
        acc >>= k;
        bits -= k;
    }
    uint32_t get(int k) {
        if (k == 0) return 0;
        const uint32_t v = peek(k);
        drop(k);
        return v;
    }
    void alignByte() { drop(bits & 7); }
};

// Canonical Huffman table (RFC 1951 3.2.2) as a lookup over the next maxLen
// bits of the stream (LSB first): entry = symbol << 4 | code length.
struct Huffman {
    std::vector<uint32_t> table;
    int maxLen = 0;
    bool build(const uint8_t* lens, int count) {
        int blCount[16] = {};
        maxLen = 0;
        for (int i = 0; i < count; ++i) {
            if (lens[i] > 15) return false;
            ++blCount[lens[i]];
            if (lens[i] > maxLen) maxLen = lens[i];
        }
        blCount[0] = 0;
        if (maxLen == 0) {   // no codes: any lookup is an error
            table.assign(1, 0);
            return true;
        }
        int next[16] = {};
        int code = 0;
        for (int b = 1; b <= 15; ++b) {
            code = (code + blCount[b - 1]) << 1;
            next[b] = code;
        }
        // over-subscribed code sets are invalid
        int left = 1;
        for (int b = 1; b <= 15; ++b) {
            left = (left << 1) - blCount[b];
            if (left < 0) return false;
        }
        table.assign((size_t)1 << maxLen, 0);
        for (int s = 0; s < count; ++s) {
            const int l = lens[s];
            if (!l) continue;
            const int c = next[l]++;
            int r = 0;
            for (int i = 0; i < l; ++i) r |= ((c >> i) & 1) << (l - 1 - i);
            for (int k = r; k < (1 << maxLen); k += 1 << l) table[(size_t)k] = (uint32_t)s << 4 | (uint32_t)l;
        }
        return true;
    }
    int decode(BitReader& br) const {
        const uint32_t e = table[br.peek(maxLen)];
        const int l = (int)(e & 15u);
        if (!l) return -1;
        br.drop(l);
        return (int)(e >> 4);
    }
};

const int kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115,
                          131, 163, 195, 227, 258};
const int kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const int kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025, 1537,
                           2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const int kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12,
                            13, 13};

bool inflateBlock(BitReader& br, const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out,
                  std::string& err) {
    for (;;) {
        const int sym = lit.decode(br);
        if (sym < 0 || br.overrun) {
            err = "corrupt DEFLATE data (bad literal/length code)";
            return false;
        }
        if (sym < 256) {
            out.push_back((uint8_t)sym);
            continue;
        }
        if (sym == 256) return true;
        const int li = sym - 257;
        if (li >= 29) {
            err = "corrupt DEFLATE data (length symbol)";
            return false;
        }
        const int len = kLenBase[li] + (int)br.get(kLenExtra[li]);
        const int ds = dist.decode(br);
        if (ds < 0 || ds >= 30) {
            err = "corrupt DEFLATE data (distance code)";
            return false;
        }
        const size_t d = (size_t)kDistBase[ds] + br.get(kDistExtra[ds]);
        if (d > out.size()) {
            err = "corrupt DEFLATE data (distance beyond the output)";
            return false;
        }
        const size_t from = out.size() - d;
        for (int i = 0; i < len; ++i) out.push_back(out[from + (size_t)i]);
    }
}

}  // namespace

bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>& out, std::string& err) {
    out.clear();
    if (size < 2) {
        err = "truncated zlib stream";
        return false;
    }
    const int cmf = data[0], flg = data[1];
    if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 != 0 || (flg & 32)) {
        err = "not a zlib DEFLATE stream (method, check bits or preset dictionary)";
        return false;
    }
    BitReader br(data + 2, size - 2);
    Huffman lit, dist;
    for (;;) {
        const uint32_t final = br.get(1);
        const uint32_t type = br.get(2);
        if (type == 0) {
            br.alignByte();
            const uint32_t len = br.get(16), nlen = br.get(16);
            if ((len ^ 0xFFFFu) != nlen) {
                err = "corrupt DEFLATE stored block";
                return false;
            }
            for (uint32_t i = 0; i < len; ++i) out.push_back((uint8_t)br.get(8));
        } else if (type == 1) {
            uint8_t lens[288];
            for (int i = 0; i < 288; ++i) lens[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
            uint8_t dl[30];
            std::memset(dl, 5, sizeof dl);
            lit.build(lens, 288);
            dist.build(dl, 30);
            if (!inflateBlock(br, lit, dist, out, err)) return false;
        } else if (type == 2) {
            const int hlit = (int)br.get(5) + 257, hdist = (int)br.get(5) + 1, hclen = (int)br.get(4) + 4;
            static const int order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {};
            for (int i = 0; i < hclen; ++i) cl[order[i]] = (uint8_t)br.get(3);
            Huffman clh;
            if (!clh.build(cl, 19)) {
                err = "corrupt DEFLATE code-length code";
                return false;
            }
            uint8_t lens[320] = {};
            int k = 0;
            while (k < hlit + hdist) {
                const int s = clh.decode(br);
                if (s < 0 || br.overrun) {
                    err = "corrupt DEFLATE code lengths";
                    return false;
                }
                if (s < 16) {
                    lens[k++] = (uint8_t)s;
                    continue;
                }
                int rep = 0;
                uint8_t v = 0;
                if (s == 16) {
                    if (k == 0) {
                        err = "corrupt DEFLATE code lengths (repeat with no previous length)";
                        return false;
                    }
                    v = lens[k - 1];
                    rep = 3 + (int)br.get(2);
                } else if (s == 17) {
                    rep = 3 + (int)br.get(3);
                } else {
                    rep = 11 + (int)br.get(7);
                }
                if (k + rep > hlit + hdist) {
                    err = "corrupt DEFLATE code lengths (repeat past the end)";
                    return false;
                }
                while (rep--) lens[k++] = v;
            }
            if (!lit.build(lens, hlit) || !dist.build(lens + hlit, hdist)) {
                err = "corrupt DEFLATE Huffman code";
                return false;
            }
            if (!inflateBlock(br, lit, dist, out, err)) return false;
        } else {
            err = "corrupt DEFLATE block type";
            return false;
        }
        if (br.overrun) {
            err = "truncated DEFLATE stream";
            return false;
        }
        if (final) return true;
    }
}

namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undoes the scanline filters of a w x h (sub)image in place; returns false
// on an unknown filter type.  `src` holds h rows of 1 + rowBytes bytes.
bool unfilter(const uint8_t* src, int h, size_t rowBytes, int bpp, std::vector<uint8_t>& dst) {
    dst.assign((size_t)h * rowBytes, 0);
    for (int y = 0; y < h; ++y) {
        const uint8_t* in = src + (size_t)y * (rowBytes + 1);
        const int type = in[0];
        ++in;
        uint8_t* row = dst.data() + (size_t)y * rowBytes;
        const uint8_t* up = y ? row - rowBytes : nullptr;
        for (size_t i = 0; i < rowBytes; ++i) {
            const int a = i >= (size_t)bpp ? row[i - (size_t)bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= (size_t)bpp) ? up[i - (size_t)bpp] : 0;
            int v;
            switch (type) {
                case 0: v = in[i]; break;
                case 1: v = in[i] + a; break;
                case 2: v = in[i] + b; break;
                case 3: v = in[i] + ((a + b) >> 1); break;
                case 4: v = in[i] + paeth(a, b, c); break;
                default: return false;
            }
            row[i] = (uint8_t)v;
        }
    }
    return true;
}

}  // namespace

bool decode_png(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                std::string& err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size < 8 || std::memcmp(data, sig, 8) != 0) {
        err = "not a PNG file (no signature)";
        return false;
    }
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    bool haveHeader = false;
    std::vector<uint8_t> idat, plte;
    std::vector<uint8_t> trns;
    size_t pos = 8;
    bool ended = false;
    while (pos + 8 <= size && !ended) {
        const uint32_t len = be32(data + pos);
        const uint8_t* type = data + pos + 4;
        if (len > size - pos - 8 || size - pos - 8 - len < 4) {
            err = "truncated PNG chunk";
            return false;
        }
        const uint8_t* body = data + pos + 8;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) {
                err = "bad IHDR chunk";
                return false;
            }
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            if (body[10] != 0 || body[11] != 0 || interlace > 1) {
                err = "unsupported PNG compression, filter or interlace method";
                return false;
            }
            haveHeader = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            ended = true;
        } else if (!(type[0] & 32)) {
            err = std::string("unknown critical PNG chunk ") + std::string((const char*)type, 4);
            return false;
        }
        pos += 12 + (size_t)len;
    }
    if (!haveHeader || idat.empty()) {
        err = "PNG without IHDR or image data";
        return false;
    }
    const bool okDepth =
        (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
        (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
        ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!okDepth) {
        err = "invalid PNG colour type / bit depth";
        return false;
    }
    if (w == 0 || h == 0 || w > 65536 || h > 65536) {
        err = "unsupported PNG size";
        return false;
    }
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) {
        err = "palette PNG without a valid PLTE chunk";
        return false;
    }
    std::vector<uint8_t> raw;
    if (!zlib_inflate(idat.data(), idat.size(), raw, err)) return false;

    const int channels = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const int bitsPP = channels * depth;
    const int bpp = bitsPP >= 8 ? bitsPP / 8 : 1;
    width = (int)w;
    height = (int)h;
    rgba.assign((size_t)w * h * 4, 0);

    // colour key (tRNS for grey / RGB), in sample units before any scaling
    int key[3] = {-1, -1, -1};
    if (ctype == 0 && trns.size() >= 2) key[0] = (trns[0] << 8 | trns[1]);
    if (ctype == 2 && trns.size() >= 6)
        for (int k = 0; k < 3; ++k) key[k] = (trns[2 * k] << 8 | trns[2 * k + 1]);
    const int greyScale = depth == 1 ? 255 : depth == 2 ? 85 : depth == 4 ? 17 : 1;

    auto sample = [&](const uint8_t* row, uint32_t x, int c) -> int {
        if (depth == 8) return row[(size_t)x * channels + c];
        if (depth == 16) return row[((size_t)x * channels + c) * 2] << 8 | row[((size_t)x * channels + c) * 2 + 1];
        const size_t bit = (size_t)x * depth;   // sub-byte depths have one channel
        return (row[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1 << depth) - 1);
    };
    auto put = [&](const uint8_t* row, uint32_t x, uint32_t ox, uint32_t oy) {
        uint8_t* o = rgba.data() + ((size_t)oy * w + ox) * 4;
        auto to8 = [&](int v) { return (uint8_t)(depth == 16 ? v >> 8 : v); };
        switch (ctype) {
            case 0: {
                const int v = sample(row, x, 0);
                const uint8_t g = depth < 8 ? (uint8_t)(v * greyScale) : to8(v);
                o[0] = o[1] = o[2] = g;
                o[3] = (key[0] >= 0 && v == key[0]) ? 0 : 255;
                break;
            }
            case 2: {
                const int r = sample(row, x, 0), g = sample(row, x, 1), b = sample(row, x, 2);
                o[0] = to8(r);
                o[1] = to8(g);
                o[2] = to8(b);
                o[3] = (key[0] >= 0 && r == key[0] && g == key[1] && b == key[2]) ? 0 : 255;
                break;
            }
            case 3: {
                const int i = sample(row, x, 0);
                if ((size_t)i * 3 + 2 < plte.size()) {
                    o[0] = plte[(size_t)i * 3];
                    o[1] = plte[(size_t)i * 3 + 1];
                    o[2] = plte[(size_t)i * 3 + 2];
                } else {
                    o[0] = o[1] = o[2] = 0;
                }
                o[3] = (size_t)i < trns.size() ? trns[(size_t)i] : 255;
                break;
            }
            case 4:
                o[0] = o[1] = o[2] = to8(sample(row, x, 0));
                o[3] = to8(sample(row, x, 1));
                break;
            default:
                for (int c = 0; c < 4; ++c) o[c] = to8(sample(row, x, c));
                break;
        }
    };

    // Adam7 passes (origin and step per pass); a non-interlaced image is one pass
    static const int a7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                 {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int passes = interlace ? 7 : 1;
    size_t off = 0;
    std::vector<uint8_t> img;
    for (int p = 0; p < passes; ++p) {
        const uint32_t x0 = interlace ? a7[p][0] : 0, y0 = interlace ? a7[p][1] : 0;
        const uint32_t dx = interlace ? a7[p][2] : 1, dy = interlace ? a7[p][3] : 1;
        const uint32_t pw = w > x0 ? (w - x0 + dx - 1) / dx : 0, ph = h > y0 ? (h - y0 + dy - 1) / dy : 0;
        if (pw == 0 || ph == 0) continue;
        const size_t rowBytes = ((size_t)pw * bitsPP + 7) / 8;
        const size_t need = (size_t)ph * (rowBytes + 1);
        if (raw.size() - off < need) {
            err = "PNG image data shorter than the image";
            return false;
        }
        if (!unfilter(raw.data() + off, (int)ph, rowBytes, bpp, img)) {
            err = "unknown PNG filter type";
            return false;
        }
        off += need;
        for (uint32_t y = 0; y < ph; ++y)
            for (uint32_t x = 0; x < pw; ++x) put(img.data() + (size_t)y * rowBytes, x, x0 + x * dx, y0 + y * dy);
    }
    return true;
}

bool decode_image(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                  std::string& err) {
    if (size >= 8 && data[0] == 137 && data[1] == 'P' && data[2] == 'N' && data[3] == 'G')
        return decode_png(data, size, width, height, rgba, err);
    if (size >= 2 && data[0] == 0xFF && data[1] == 0xD8) return decode_jpeg(data, size, width, height, rgba, err);
    err = "unsupported image format (PNG and baseline JPEG are decoded)";
    return false;
}

}  // namespace bte
