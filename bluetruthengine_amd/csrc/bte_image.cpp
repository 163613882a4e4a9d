// bte_image.cpp — baseline JPEG decoding for texture maps.
//
// The reference loads texture files with stb_image (src/main.cpp:442-475).
// This decoder follows ITU-T T.81 itself: marker segments (Annex B), Huffman
// tables and the DECODE / RECEIVE / EXTEND procedures (Annex C, F.2.2),
// dequantization, the inverse DCT of A.3.3 evaluated in binary64 and rounded,
// the level shift, and the JFIF YCbCr -> RGB conversion.  Chroma planes
// sampled at half resolution are upsampled the way stb_image does it
// (stb_image.h:3465-3527, chosen at :3909-3918, rows at :3927-3938): a triangle filter
// weighting the nearer sample 3/4 and the farther 1/4, rounded as
// (3 near + far + 2) >> 2 along one axis and (3 t0 + t1 + 8) >> 4 over the
// vertical sums t of both, edges replicated; other factors repeat the sample.
// Decoded pixels can differ by a unit from a decoder that uses a fixed-point
// IDCT.
#include "bte_image.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace bte {
namespace {

// Zig-zag sequence (T.81 figure A.6): coefficient k of the stream is element
// kZigZag[k] of the 8 x 8 block in row-major order.
const int kZigZag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    bool present = false;
    int mincode[17] = {};
    int maxcode[18] = {};   // -1: no code of that length
    int valptr[17] = {};
    uint8_t vals[256] = {};
};

// Canonical code assignment (T.81 C.2): codes of each length follow the
// previous length's last code, shifted left by one.
bool buildHuff(Huff& h, const uint8_t* counts, const uint8_t* syms, int nsyms) {
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        const int n = counts[l - 1];
        h.valptr[l] = k;
        h.mincode[l] = code;
        code += n;
        k += n;
        h.maxcode[l] = n ? code - 1 : -1;
        if (code > (1 << l)) return false;
        code <<= 1;
    }
    h.maxcode[17] = INT_MAX;
    if (k != nsyms || k > 256) return false;
    std::memcpy(h.vals, syms, (size_t)nsyms);
    h.present = true;
    return true;
}

// Entropy-coded data, most significant bit first; 0xFF00 is a data byte 0xFF,
// any other 0xFF xx is a marker, after which zeros are read.
struct BitReader {
    const uint8_t* p;
    const uint8_t* end;
    uint32_t acc = 0;
    int nbits = 0;
    bool atMarker = false;

    int bit() {
        if (nbits == 0) {
            int byte = 0;
            if (!atMarker && p < end) {
                byte = *p;
                if (byte == 0xFF) {
                    const int next = p + 1 < end ? p[1] : 0xD9;
                    if (next == 0x00) {
                        p += 2;
                    } else {
                        atMarker = true;
                        byte = 0;
                    }
                } else {
                    ++p;
                }
            }
            acc = (uint32_t)byte;
            nbits = 8;
        }
        --nbits;
        return (int)((acc >> nbits) & 1u);
    }
    int receive(int s) {
        int v = 0;
        for (int i = 0; i < s; ++i) v = (v << 1) | bit();
        return v;
    }
    // Restart interval boundary: drop the rest of the byte and the RSTn marker.
    void restart() {
        nbits = 0;
        atMarker = false;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) ++p;
        if (p + 1 < end) p += 2;
    }
};

int decodeSymbol(BitReader& br, const Huff& h, bool& ok) {
    int code = br.bit();
    int l = 1;
    while (code > h.maxcode[l]) {
        code = (code << 1) | br.bit();
        if (++l > 16) {
            ok = false;
            return 0;
        }
    }
    return h.vals[h.valptr[l] + code - h.mincode[l]];
}

int extendValue(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;       // Huffman tables of the current scan
    int pred = 0;             // DC predictor
    int planeW = 0, planeH = 0;
    std::vector<uint8_t> plane;
};

// Produces a component's samples at full resolution, one output row at a
// time.  hs x vs = hmax / h x vmax / v; the component holds ceil(W / hs)
// x ceil(H v / vmax) samples.  The two source rows of output row y: the
// nearer one, row floor(y / vs), and the farther one, the row above for the
// upper half of a pair and the row below for the lower half (clamped to the
// plane), advanced as a two-row window per vs output rows.
struct Upsampler {
    int hs = 1, vs = 1, wLo = 0, rowsLo = 0;
    int ystep = 0, ypos = 0, line0 = 0, line1 = 0;
    std::vector<uint8_t> row;
    void init(const Component& c, int W, int H, int hmax, int vmax) {
        hs = std::max(1, hmax / c.h);
        vs = std::max(1, vmax / c.v);
        wLo = (W + hs - 1) / hs;
        rowsLo = (H * c.v + vmax - 1) / vmax;
        ystep = vs >> 1;
        ypos = 0;
        line0 = line1 = 0;
        row.assign((size_t)wLo * hs + 1, 0);
    }
    void nextRow(const Component& c) {
        const bool bottom = ystep >= (vs >> 1);
        const uint8_t* nearRow = &c.plane[(size_t)(bottom ? line1 : line0) * c.planeW];
        const uint8_t* farRow = &c.plane[(size_t)(bottom ? line0 : line1) * c.planeW];
        uint8_t* out = row.data();
        if (hs == 1 && vs == 1) {
            std::memcpy(out, nearRow, (size_t)wLo);
        } else if (hs == 1 && vs == 2) {
            for (int i = 0; i < wLo; ++i) out[i] = (uint8_t)((3 * nearRow[i] + farRow[i] + 2) >> 2);
        } else if (hs == 2 && vs == 1) {
            if (wLo == 1) {
                out[0] = out[1] = nearRow[0];
            } else {
                out[0] = nearRow[0];
                out[1] = (uint8_t)((nearRow[0] * 3 + nearRow[1] + 2) >> 2);
                for (int i = 1; i < wLo - 1; ++i) {
                    const int n = 3 * nearRow[i] + 2;
                    out[2 * i] = (uint8_t)((n + nearRow[i - 1]) >> 2);
                    out[2 * i + 1] = (uint8_t)((n + nearRow[i + 1]) >> 2);
                }
                out[2 * (wLo - 1)] = (uint8_t)((nearRow[wLo - 2] * 3 + nearRow[wLo - 1] + 2) >> 2);
                out[2 * (wLo - 1) + 1] = nearRow[wLo - 1];
            }
        } else if (hs == 2 && vs == 2) {
            int t1 = 3 * nearRow[0] + farRow[0];
            if (wLo == 1) {
                out[0] = out[1] = (uint8_t)((t1 + 2) >> 2);
            } else {
                out[0] = (uint8_t)((t1 + 2) >> 2);
                for (int i = 1; i < wLo; ++i) {
                    const int t0 = t1;
                    t1 = 3 * nearRow[i] + farRow[i];
                    out[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
                    out[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
                }
                out[2 * wLo - 1] = (uint8_t)((t1 + 2) >> 2);
            }
        } else {
            for (int i = 0; i < wLo; ++i)
                for (int j = 0; j < hs; ++j) out[(size_t)i * hs + j] = nearRow[i];
        }
        if (++ystep >= vs) {
            ystep = 0;
            line0 = line1;
            if (++ypos < rowsLo) ++line1;
        }
    }
};

// kIdct[x][u] = C(u) / 2 * cos((2x + 1) u pi / 16), C(0) = 1 / sqrt(2).
struct IdctTable {
    double c[8][8];
    IdctTable() {
        const double pi = std::acos(-1.0);
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                c[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * pi / 16.0);
    }
};
const IdctTable kIdct;

void idctBlock(const int* coef, const uint16_t* q, uint8_t* out, int stride) {
    double f[64], t[64];
    for (int i = 0; i < 64; ++i) f[i] = (double)coef[i] * (double)q[i];
    for (int v = 0; v < 8; ++v)
        for (int x = 0; x < 8; ++x) {
            double s = 0.0;
            for (int u = 0; u < 8; ++u) s += kIdct.c[x][u] * f[v * 8 + u];
            t[v * 8 + x] = s;
        }
    for (int y = 0; y < 8; ++y)
        for (int x = 0; x < 8; ++x) {
            double s = 0.0;
            for (int v = 0; v < 8; ++v) s += kIdct.c[y][v] * t[v * 8 + x];
            const double val = std::floor(s + 128.0 + 0.5);
            out[y * stride + x] = (uint8_t)(val < 0.0 ? 0.0 : (val > 255.0 ? 255.0 : val));
        }
}

uint8_t clampByte(double x) {
    const double r = std::floor(x + 0.5);
    return (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
}

}  // namespace

bool decode_jpeg(const uint8_t* data, size_t size, int& width, int& height, std::vector<uint8_t>& rgba,
                 std::string& err) {
    auto bad = [&](const char* why) {
        err = why;
        return false;
    };
    if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) return bad("not a JPEG file (no SOI marker)");
    uint16_t qt[4][64];
    bool qtPresent[4] = {false, false, false, false};
    Huff dc[4], ac[4];
    std::vector<Component> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, restartInterval = 0;
    bool frameSeen = false, adobeRgb = false, scanned = false;
    size_t pos = 2;
    while (pos + 4 <= size) {
        if (data[pos] != 0xFF) return bad("corrupt marker sequence");
        const int marker = data[pos + 1];
        if (marker == 0xFF) {      // fill byte
            ++pos;
            continue;
        }
        pos += 2;
        if (marker == 0xD9) break;                              // EOI
        if (marker == 0x01 || (marker >= 0xD0 && marker <= 0xD7)) continue;
        if (pos + 2 > size) return bad("truncated marker segment");
        const size_t len = ((size_t)data[pos] << 8) | data[pos + 1];
        if (len < 2 || pos + len > size) return bad("truncated marker segment");
        const uint8_t* seg = data + pos + 2;
        const size_t n = len - 2;
        if (marker == 0xDB) {                                   // DQT
            size_t i = 0;
            while (i < n) {
                const int pq = seg[i] >> 4, tq = seg[i] & 15;
                ++i;
                if (tq > 3 || i + (pq ? 128 : 64) > n) return bad("bad quantization table");
                for (int k = 0; k < 64; ++k) {
                    const uint16_t val = pq ? (uint16_t)((seg[i + 2 * k] << 8) | seg[i + 2 * k + 1]) : seg[i + k];
                    qt[tq][kZigZag[k]] = val;
                }
                qtPresent[tq] = true;
                i += pq ? 128 : 64;
            }
        } else if (marker == 0xC4) {                            // DHT
            size_t i = 0;
            while (i < n) {
                if (i + 17 > n) return bad("bad Huffman table");
                const int tc = seg[i] >> 4, th = seg[i] & 15;
                int total = 0;
                for (int k = 0; k < 16; ++k) total += seg[i + 1 + k];
                if (tc > 1 || th > 3 || i + 17 + (size_t)total > n) return bad("bad Huffman table");
                Huff& h = tc == 0 ? dc[th] : ac[th];
                if (!buildHuff(h, seg + i + 1, seg + i + 17, total)) return bad("bad Huffman table");
                i += 17 + (size_t)total;
            }
        } else if (marker == 0xC0 || marker == 0xC1) {          // SOF0 / SOF1: sequential, Huffman
            if (n < 6 || seg[0] != 8) return bad("only 8-bit sequential JPEG is supported");
            H = (seg[1] << 8) | seg[2];
            W = (seg[3] << 8) | seg[4];
            const int nc = seg[5];
            if (W <= 0 || H <= 0 || (nc != 1 && nc != 3) || n < 6 + 3 * (size_t)nc)
                return bad("unsupported frame (size or component count)");
            comps.assign((size_t)nc, Component{});
            for (int c = 0; c < nc; ++c) {
                comps[c].id = seg[6 + 3 * c];
                comps[c].h = seg[7 + 3 * c] >> 4;
                comps[c].v = seg[7 + 3 * c] & 15;
                comps[c].tq = seg[8 + 3 * c];
                if (comps[c].h < 1 || comps[c].h > 4 || comps[c].v < 1 || comps[c].v > 4 || comps[c].tq > 3)
                    return bad("bad component sampling");
                hmax = std::max(hmax, comps[c].h);
                vmax = std::max(vmax, comps[c].v);
            }
            const int mcuCols = (W + 8 * hmax - 1) / (8 * hmax), mcuRows = (H + 8 * vmax - 1) / (8 * vmax);
            for (Component& c : comps) {
                c.planeW = mcuCols * c.h * 8;
                c.planeH = mcuRows * c.v * 8;
                c.plane.assign((size_t)c.planeW * (size_t)c.planeH, 0);
            }
            frameSeen = true;
        } else if ((marker >= 0xC2 && marker <= 0xC3) || (marker >= 0xC5 && marker <= 0xCF && marker != 0xC8 &&
                                                          marker != 0xCC)) {
            return bad("progressive, lossless and arithmetic-coded JPEG are not supported");
        } else if (marker == 0xDD) {                            // DRI
            if (n < 2) return bad("bad restart interval");
            restartInterval = (seg[0] << 8) | seg[1];
        } else if (marker == 0xEE) {                            // APP14: Adobe colour transform flag
            if (n >= 12 && std::memcmp(seg, "Adobe", 5) == 0 && seg[11] == 0 && comps.size() != 1) adobeRgb = true;
        } else if (marker == 0xDA) {                            // SOS
            if (!frameSeen) return bad("scan before frame header");
            const int ns = n > 0 ? seg[0] : 0;
            if (ns < 1 || ns > 4 || n < 4 + 2 * (size_t)ns) return bad("bad scan header");
            std::vector<Component*> sc;
            for (int k = 0; k < ns; ++k) {
                Component* c = nullptr;
                for (Component& x : comps)
                    if (x.id == seg[1 + 2 * k]) c = &x;
                if (!c) return bad("scan names an unknown component");
                c->td = seg[2 + 2 * k] >> 4;
                c->ta = seg[2 + 2 * k] & 15;
                if (c->td > 3 || c->ta > 3 || !dc[c->td].present || !ac[c->ta].present || !qtPresent[c->tq])
                    return bad("scan refers to a missing table");
                c->pred = 0;
                sc.push_back(c);
            }
            const bool interleaved = ns > 1;
            int mcuCols, mcuRows;
            if (interleaved) {
                mcuCols = (W + 8 * hmax - 1) / (8 * hmax);
                mcuRows = (H + 8 * vmax - 1) / (8 * vmax);
            } else {
                const int cw = (W * sc[0]->h + hmax - 1) / hmax, ch = (H * sc[0]->v + vmax - 1) / vmax;
                mcuCols = (cw + 7) / 8;
                mcuRows = (ch + 7) / 8;
            }
            BitReader br{data + pos + len, data + size};
            int coef[64];
            bool ok = true;
            long mcu = 0;
            for (int my = 0; my < mcuRows && ok; ++my)
                for (int mx = 0; mx < mcuCols && ok; ++mx, ++mcu) {
                    if (restartInterval > 0 && mcu > 0 && mcu % restartInterval == 0) {
                        br.restart();
                        for (Component* c : sc) c->pred = 0;
                    }
                    for (Component* c : sc) {
                        const int bh = interleaved ? c->h : 1, bv = interleaved ? c->v : 1;
                        for (int v = 0; v < bv && ok; ++v)
                            for (int h = 0; h < bh && ok; ++h) {
                                std::memset(coef, 0, sizeof coef);
                                const int t = decodeSymbol(br, dc[c->td], ok);
                                if (!ok || t > 11) {
                                    ok = false;
                                    break;
                                }
                                c->pred += extendValue(br.receive(t), t);
                                coef[0] = c->pred;
                                for (int k = 1; k < 64;) {
                                    const int rs = decodeSymbol(br, ac[c->ta], ok);
                                    if (!ok) break;
                                    const int r = rs >> 4, s = rs & 15;
                                    if (s == 0) {
                                        if (r != 15) break;   // end of block
                                        k += 16;
                                        continue;
                                    }
                                    k += r;
                                    if (k > 63) {
                                        ok = false;
                                        break;
                                    }
                                    coef[kZigZag[k]] = extendValue(br.receive(s), s);
                                    ++k;
                                }
                                const int bx = interleaved ? mx * c->h + h : mx, by = interleaved ? my * c->v + v : my;
                                if ((bx + 1) * 8 <= c->planeW && (by + 1) * 8 <= c->planeH)
                                    idctBlock(coef, qt[c->tq], &c->plane[(size_t)by * 8 * c->planeW + (size_t)bx * 8],
                                              c->planeW);
                            }
                    }
                }
            if (!ok) return bad("corrupt entropy-coded data");
            scanned = true;
            // continue after the entropy-coded segment: the next marker
            size_t q = (size_t)(br.p - data);
            while (q + 1 < size && !(data[q] == 0xFF && data[q + 1] != 0x00 && !(data[q + 1] >= 0xD0 && data[q + 1] <= 0xD7)))
                ++q;
            pos = q;
            continue;
        }
        pos += len;
    }
    if (!frameSeen || !scanned) return bad("no image data");
    width = W;
    height = H;
    rgba.assign((size_t)W * (size_t)H * 4, 255);
    // per component: its full-resolution row for the current output row
    std::vector<Upsampler> up(comps.size());
    for (size_t c = 0; c < comps.size(); ++c) up[c].init(comps[c], W, H, hmax, vmax);
    for (int y = 0; y < H; ++y) {
        for (size_t c = 0; c < comps.size(); ++c) up[c].nextRow(comps[c]);
        for (int x = 0; x < W; ++x) {
            uint8_t* o = &rgba[((size_t)y * W + x) * 4];
            double s[3] = {0.0, 0.0, 0.0};
            for (size_t c = 0; c < comps.size(); ++c) s[c] = up[c].row[(size_t)x];
            if (comps.size() == 1) {
                o[0] = o[1] = o[2] = (uint8_t)s[0];
            } else if (adobeRgb) {
                o[0] = (uint8_t)s[0];
                o[1] = (uint8_t)s[1];
                o[2] = (uint8_t)s[2];
            } else {   // JFIF YCbCr
                const double yy = s[0], cb = s[1] - 128.0, cr = s[2] - 128.0;
                o[0] = clampByte(yy + 1.402 * cr);
                o[1] = clampByte(yy - 0.344136 * cb - 0.714136 * cr);
                o[2] = clampByte(yy + 1.772 * cb);
            }
        }
    }
    return true;
}

}  // namespace bte
