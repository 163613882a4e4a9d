// bte_gltf.cpp — glTF 2.0 / GLB ingestion for <model> elements.
//
// Follows src/gltf_loader.cpp, which reads the file with the vendored
// tinygltf and flattens it with GLM matrices; here the container, the JSON
// (bte_json.cpp) and the images (bte_image.cpp / bte_png.cpp) are read by
// this project's own code and the matrices are plain binary32 4 x 4 arrays:
//   placement matrix                xmlTransformToMat4 :40-50 (T Rz Ry Rx S)
//   node matrices                   nodeLocalMatrix :52-85 (matrix, else T R S)
//   accessors                       readFloatComponent :101-144 (normalized
//                                   integers), readIndexComponent :146-163,
//                                   accessorPointer :165-195, readVec3Attr /
//                                   readVec2Attr / readIndices :197-259
//   primitive modes                 triangulate :261-293 (lists, fans, strips)
//   materials                       disneyFromGltf :369-454 (KHR_materials_ior,
//                                   _transmission, _clearcoat, _sheen,
//                                   _emissive_strength, _unlit), texture
//                                   factor baking :351-367, defaultMaterial :456-467
//   triangles                       emitPrimitive :490-580 (normal matrix,
//                                   geometric-normal fallback, TEXCOORD_0)
//   skipped parts                   skipBackdropNode :589-594,
//                                   skipToonOrBackdropPrim :596-603,
//                                   outlierPrims :689-727 (extent > 12x or
//                                   centre farther than 10x the median extent)
//   traversal and clean-up          walkNode :729-766, loadGltfModel :768-846
// Image decoding follows tinygltf's stb_image path with 4 requested channels.
#include "bte_gltf.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <set>

#include "bte_image.h"
#include "bte_json.h"

namespace bte {
namespace {

struct M4 {
    float m[4][4];   // row-major: p' = M p for column vectors
};

M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0f;
    return r;
}

M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = a.m[i][0] * b.m[0][j];
            for (int k = 1; k < 4; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}

M4 translation(float x, float y, float z) {
    M4 r = identity();
    r.m[0][3] = x;
    r.m[1][3] = y;
    r.m[2][3] = z;
    return r;
}

M4 scaling(float x, float y, float z) {
    M4 r = identity();
    r.m[0][0] = x;
    r.m[1][1] = y;
    r.m[2][2] = z;
    return r;
}

// right-handed rotation by `rad` about coordinate axis 0 (x), 1 (y) or 2 (z)
M4 rotation(int axis, float rad) {
    const float c = std::cos(rad), s = std::sin(rad);
    M4 r = identity();
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    r.m[a][a] = c;
    r.m[a][b] = -s;
    r.m[b][a] = s;
    r.m[b][b] = c;
    return r;
}

// rotation matrix of the quaternion (x, y, z, w), as given (not renormalised)
M4 quaternion(float x, float y, float z, float w) {
    M4 r = identity();
    r.m[0][0] = 1.0f - 2.0f * (y * y + z * z);
    r.m[0][1] = 2.0f * (x * y - w * z);
    r.m[0][2] = 2.0f * (x * z + w * y);
    r.m[1][0] = 2.0f * (x * y + w * z);
    r.m[1][1] = 1.0f - 2.0f * (x * x + z * z);
    r.m[1][2] = 2.0f * (y * z - w * x);
    r.m[2][0] = 2.0f * (x * z - w * y);
    r.m[2][1] = 2.0f * (y * z + w * x);
    r.m[2][2] = 1.0f - 2.0f * (x * x + y * y);
    return r;
}

BteFloat3 xformPoint(const M4& M, float x, float y, float z) {
    BteFloat3 o;
    o.x = (M.m[0][0] * x + M.m[0][1] * y) + (M.m[0][2] * z + M.m[0][3]);
    o.y = (M.m[1][0] * x + M.m[1][1] * y) + (M.m[1][2] * z + M.m[1][3]);
    o.z = (M.m[2][0] * x + M.m[2][1] * y) + (M.m[2][2] * z + M.m[2][3]);
    return o;
}

// transpose(inverse(upper 3 x 3)) = cofactor matrix / determinant
void normalMatrix(const M4& M, float N[3][3]) {
    const float(*a)[4] = M.m;
    float cof[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cof[i][j] = a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1];
        }
    const float det = a[0][0] * cof[0][0] + a[0][1] * cof[0][1] + a[0][2] * cof[0][2];
    const float inv = 1.0f / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) N[i][j] = cof[i][j] * inv;
}

BteFloat3 normalize(float x, float y, float z) {
    const float r = 1.0f / std::sqrt(x * x + y * y + z * z);
    return {x * r, y * r, z * r};
}

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool readAll(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n <= 0 || std::fread(out.data(), 1, (size_t)n, f) == (size_t)n;
    std::fclose(f);
    return ok;
}

bool base64(const std::string& s, size_t from, std::vector<uint8_t>& out) {
    out.clear();
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        const char c = s[i];
        int v;
        if (c >= 'A' && c <= 'Z') v = c - 'A';
        else if (c >= 'a' && c <= 'z') v = c - 'a' + 26;
        else if (c >= '0' && c <= '9') v = c - '0' + 52;
        else if (c == '+' || c == '-') v = 62;
        else if (c == '/' || c == '_') v = 63;
        else if (c == '=' || c == '\n' || c == '\r' || c == ' ') continue;
        else return false;
        acc = acc << 6 | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return true;
}

std::string percentDecode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
            std::isxdigit((unsigned char)s[i + 2])) {
            o += (char)std::strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            o += s[i];
        }
    }
    return o;
}

struct Doc {
    JValue j;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;   // directory of the model file (with a trailing '/'), for relative URIs
    std::string err;

    const JValue* item(const char* list, int i) const {
        const JValue* l = j.get(list);
        if (!l || !l->isArray() || i < 0 || (size_t)i >= l->arr.size()) return nullptr;
        return &l->arr[(size_t)i];
    }
    size_t count(const char* list) const {
        const JValue* l = j.get(list);
        return (l && l->isArray()) ? l->arr.size() : 0;
    }
    bool loadUri(const std::string& uri, std::vector<uint8_t>& out) {
        if (uri.compare(0, 5, "data:") == 0) {
            const size_t k = uri.find(";base64,");
            if (k == std::string::npos || !base64(uri, k + 8, out)) {
                err = "unsupported or corrupt data URI";
                return false;
            }
            return true;
        }
        const std::string path = (!uri.empty() && uri[0] == '/') ? uri : dir + percentDecode(uri);
        if (!readAll(path, out)) {
            err = "cannot read \"" + path + "\"";
            return false;
        }
        return true;
    }
};

// a non-negative whole byte offset / length / count that fits comfortably in size_t
bool toSize(double v, size_t& out) {
    if (!(v >= 0.0 && v <= 4503599627370496.0)) return false;   // also rejects NaN
    out = (size_t)v;
    return true;
}

// a JSON number used as an index: -1 (no such item) when outside int
int toIndex(double v) { return (v >= 0.0 && v <= 2147483647.0) ? (int)v : -1; }

int componentCount(const std::string& type) {
    if (type == "SCALAR") return 1;
    if (type == "VEC2") return 2;
    if (type == "VEC3") return 3;
    if (type == "VEC4") return 4;
    return 0;
}

int componentSize(int ct) {
    switch (ct) {
        case 5120: case 5121: return 1;   // BYTE, UNSIGNED_BYTE
        case 5122: case 5123: return 2;   // SHORT, UNSIGNED_SHORT
        case 5125: case 5126: return 4;   // UNSIGNED_INT, FLOAT
        default: return 0;
    }
}

float readComponent(const uint8_t* p, int ct, bool normalized) {
    switch (ct) {
        case 5120: {
            const int8_t v = (int8_t)*p;
            return normalized ? std::max((float)v / 127.0f, -1.0f) : (float)v;
        }
        case 5121: return normalized ? (float)*p / 255.0f : (float)*p;
        case 5122: {
            int16_t v;
            std::memcpy(&v, p, 2);
            return normalized ? std::max((float)v / 32767.0f, -1.0f) : (float)v;
        }
        case 5123: {
            uint16_t v;
            std::memcpy(&v, p, 2);
            return normalized ? (float)v / 65535.0f : (float)v;
        }
        case 5125: {
            uint32_t v;
            std::memcpy(&v, p, 4);
            return (float)v;
        }
        case 5126: {
            float v;
            std::memcpy(&v, p, 4);
            return v;
        }
        default: return 0.0f;
    }
}

uint32_t readIndex(const uint8_t* p, int ct) {
    switch (ct) {
        case 5121: return *p;
        case 5123: { uint16_t v; std::memcpy(&v, p, 2); return v; }
        case 5125: { uint32_t v; std::memcpy(&v, p, 4); return v; }
        default: return 0;
    }
}

struct View {
    const uint8_t* p = nullptr;
    size_t stride = 0, count = 0;
    int ct = 0, comps = 0;
    bool normalized = false;
};

// accessor -> element pointer, stride and count; false when any reference
// is missing or the last element would overrun its buffer
bool accessor(const Doc& d, int index, View& v) {
    const JValue* acc = d.item("accessors", index);
    if (!acc) return false;
    const JValue* bv = d.item("bufferViews", acc->integer("bufferView", -1));
    if (!bv) return false;
    const int b = bv->integer("buffer", -1);
    if (b < 0 || (size_t)b >= d.buffers.size()) return false;
    const std::vector<uint8_t>& buf = d.buffers[(size_t)b];
    v.comps = componentCount(acc->string("type"));
    v.ct = acc->integer("componentType", 0);
    const int cs = componentSize(v.ct);
    if (v.comps <= 0 || cs <= 0) return false;
    const size_t elem = (size_t)v.comps * (size_t)cs;
    const int byteStride = bv->integer("byteStride", 0);
    v.stride = byteStride > 0 ? (size_t)byteStride : elem;
    size_t viewOff = 0, accOff = 0;
    if (!toSize(acc->number("count", 0.0), v.count) || !toSize(bv->number("byteOffset", 0.0), viewOff) ||
        !toSize(acc->number("byteOffset", 0.0), accOff))
        return false;
    // every element occupies at least one byte of the buffer: bounds the products below
    if (v.count > buf.size() || viewOff > buf.size() || accOff > buf.size()) return false;
    const size_t offset = viewOff + accOff;
    if (offset + (v.count ? (v.count - 1) * v.stride + elem : 0) > buf.size()) return false;
    v.p = buf.data() + offset;
    const JValue* nz = acc->get("normalized");
    v.normalized = nz && nz->type == JValue::Bool && nz->b;
    return true;
}

std::vector<BteFloat3> readVec3(const Doc& d, int index) {
    std::vector<BteFloat3> out;
    View v;
    if (!accessor(d, index, v) || v.comps < 3) return out;
    const int cs = componentSize(v.ct);
    out.resize(v.count);
    for (size_t i = 0; i < v.count; ++i) {
        const uint8_t* e = v.p + i * v.stride;
        out[i] = {readComponent(e, v.ct, v.normalized), readComponent(e + cs, v.ct, v.normalized),
                  readComponent(e + 2 * cs, v.ct, v.normalized)};
    }
    return out;
}

std::vector<BteFloat2> readVec2(const Doc& d, int index) {
    std::vector<BteFloat2> out;
    View v;
    if (!accessor(d, index, v) || v.comps < 2) return out;
    const int cs = componentSize(v.ct);
    out.resize(v.count);
    for (size_t i = 0; i < v.count; ++i) {
        const uint8_t* e = v.p + i * v.stride;
        out[i] = {readComponent(e, v.ct, v.normalized), readComponent(e + cs, v.ct, v.normalized)};
    }
    return out;
}

std::vector<uint32_t> readIndices(const Doc& d, int index, size_t vertexCount) {
    std::vector<uint32_t> out;
    if (index < 0) {
        out.resize(vertexCount);
        for (size_t i = 0; i < vertexCount; ++i) out[i] = (uint32_t)i;
        return out;
    }
    View v;
    if (!accessor(d, index, v)) return out;
    out.resize(v.count);
    for (size_t i = 0; i < v.count; ++i) out[i] = readIndex(v.p + i * v.stride, v.ct);
    return out;
}

// modes 4 (list), 5 (strip: odd triangles swap their first two corners), 6 (fan)
std::vector<uint32_t> triangulate(const std::vector<uint32_t>& idx, int mode) {
    std::vector<uint32_t> t;
    if (mode == 4) {
        t.assign(idx.begin(), idx.begin() + (std::ptrdiff_t)(idx.size() - idx.size() % 3));
    } else if (mode == 6 && idx.size() >= 3) {
        for (size_t i = 1; i + 1 < idx.size(); ++i) t.insert(t.end(), {idx[0], idx[i], idx[i + 1]});
    } else if (mode == 5 && idx.size() >= 3) {
        for (size_t i = 0; i + 2 < idx.size(); ++i) {
            if (i % 2 == 0) t.insert(t.end(), {idx[i], idx[i + 1], idx[i + 2]});
            else t.insert(t.end(), {idx[i + 1], idx[i], idx[i + 2]});
        }
    }
    return t;
}

M4 nodeMatrix(const JValue& node) {
    const std::vector<double> mat = node.numbers("matrix");
    if (mat.size() == 16) {   // column-major in the file
        M4 r{};
        for (int c = 0; c < 4; ++c)
            for (int rr = 0; rr < 4; ++rr) r.m[rr][c] = (float)mat[(size_t)(c * 4 + rr)];
        return r;
    }
    M4 T = identity(), R = identity(), S = identity();
    const std::vector<double> t = node.numbers("translation"), q = node.numbers("rotation"), s = node.numbers("scale");
    if (t.size() == 3) T = translation((float)t[0], (float)t[1], (float)t[2]);
    if (q.size() == 4) R = quaternion((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
    if (s.size() == 3) S = scaling((float)s[0], (float)s[1], (float)s[2]);
    return mul(mul(T, R), S);
}

bool skipNode(const JValue& node) {
    const std::string n = lower(node.string("name"));
    return n == "background" || n.find("backgroundgradient") != std::string::npos;
}

bool skipPrim(const Doc& d, const JValue& prim) {
    const JValue* m = d.item("materials", prim.integer("material", -1));
    if (!m) return false;
    const std::string n = lower(m->string("name"));
    return n.find("outline") != std::string::npos || n == "backgroundgradient";
}

uint64_t primKey(int node, int prim) { return (uint64_t)(uint32_t)node << 32 | (uint32_t)prim; }

const JValue* attribute(const JValue& prim, const char* name) {
    const JValue* a = prim.get("attributes");
    const JValue* v = a ? a->get(name) : nullptr;
    return (v && v->isNumber()) ? v : nullptr;
}

struct PrimStat {
    uint64_t key;
    float c[3];
    float extent;
};

const int kMaxDepth = 1024;

bool gatherStats(const Doc& d, int ni, const M4& parent, std::vector<PrimStat>& stats, int depth) {
    const JValue* node = d.item("nodes", ni);
    if (!node) return true;
    if (depth > kMaxDepth) return false;
    if (skipNode(*node)) return true;
    const M4 world = mul(parent, nodeMatrix(*node));
    const JValue* mesh = d.item("meshes", node->integer("mesh", -1));
    const JValue* prims = mesh ? mesh->get("primitives") : nullptr;
    if (prims && prims->isArray()) {
        for (size_t pi = 0; pi < prims->arr.size(); ++pi) {
            const JValue& prim = prims->arr[pi];
            if (skipPrim(d, prim)) continue;
            const JValue* pos = attribute(prim, "POSITION");
            if (!pos) continue;
            const JValue* acc = d.item("accessors", toIndex(pos->num));
            if (!acc) continue;
            const std::vector<double> lo = acc->numbers("min"), hi = acc->numbers("max");
            if (lo.size() < 3 || hi.size() < 3) continue;
            float wmin[3] = {1e30f, 1e30f, 1e30f}, wmax[3] = {-1e30f, -1e30f, -1e30f};
            for (int k = 0; k < 8; ++k) {
                const BteFloat3 w = xformPoint(world, (float)((k & 1) ? hi[0] : lo[0]), (float)((k & 2) ? hi[1] : lo[1]),
                                               (float)((k & 4) ? hi[2] : lo[2]));
                const float wv[3] = {w.x, w.y, w.z};
                for (int a = 0; a < 3; ++a) {
                    wmin[a] = std::min(wmin[a], wv[a]);
                    wmax[a] = std::max(wmax[a], wv[a]);
                }
            }
            PrimStat s;
            s.key = primKey(ni, (int)pi);
            for (int a = 0; a < 3; ++a) s.c[a] = 0.5f * (wmin[a] + wmax[a]);
            s.extent = std::max(wmax[0] - wmin[0], std::max(wmax[1] - wmin[1], wmax[2] - wmin[2]));
            if (s.extent > 0.0f) stats.push_back(s);
        }
    }
    const JValue* ch = node->get("children");
    if (ch && ch->isArray())
        for (const JValue& c : ch->arr)
            if (c.isNumber() && !gatherStats(d, toIndex(c.num), world, stats, depth + 1)) return false;
    return true;
}

// primitives far larger than, or far from, the median one (detached capes,
// colliders): extent > 12 x median extent or centre > 10 x median extent
// from the per-axis median centre; only with 4 or more primitives
std::set<uint64_t> outliers(const std::vector<PrimStat>& st) {
    std::set<uint64_t> skip;
    if (st.size() < 4) return skip;
    auto median = [](std::vector<float> v) {
        const size_t n = v.size();
        std::nth_element(v.begin(), v.begin() + (std::ptrdiff_t)(n / 2), v.end());
        return v[n / 2];
    };
    std::vector<float> e, cx, cy, cz;
    for (const PrimStat& s : st) {
        e.push_back(s.extent);
        cx.push_back(s.c[0]);
        cy.push_back(s.c[1]);
        cz.push_back(s.c[2]);
    }
    const float me = median(e);
    const float mc[3] = {median(cx), median(cy), median(cz)};
    if (me <= 1e-8f) return skip;
    const float maxExtent = 12.0f * me, maxDist = 10.0f * me;
    for (const PrimStat& s : st) {
        const float dx = s.c[0] - mc[0], dy = s.c[1] - mc[1], dz = s.c[2] - mc[2];
        if (s.extent > maxExtent || std::sqrt(dx * dx + dy * dy + dz * dz) > maxDist) skip.insert(s.key);
    }
    return skip;
}

void emitPrimitive(const Doc& d, const JValue& prim, const M4& world, int slot, GltfModel& out) {
    const JValue* pa = attribute(prim, "POSITION");
    if (!pa) return;
    const std::vector<BteFloat3> pos = readVec3(d, toIndex(pa->num));
    if (pos.empty()) return;
    const JValue* na = attribute(prim, "NORMAL");
    const std::vector<BteFloat3> nrm = na ? readVec3(d, toIndex(na->num)) : std::vector<BteFloat3>();
    const JValue* ua = attribute(prim, "TEXCOORD_0");
    const std::vector<BteFloat2> uvs = ua ? readVec2(d, toIndex(ua->num)) : std::vector<BteFloat2>();
    const int mode = prim.integer("mode", 4) < 0 ? 4 : prim.integer("mode", 4);
    if (mode != 4 && mode != 5 && mode != 6) {
        std::fprintf(stderr, "[gltf] skipping a non-triangle primitive (mode %d)\n", mode);
        return;
    }
    const std::vector<uint32_t> idx = triangulate(readIndices(d, prim.integer("indices", -1), pos.size()), mode);
    float N[3][3];
    normalMatrix(world, N);
    auto xn = [&](const BteFloat3& n) {
        return normalize(N[0][0] * n.x + N[0][1] * n.y + N[0][2] * n.z, N[1][0] * n.x + N[1][1] * n.y + N[1][2] * n.z,
                         N[2][0] * n.x + N[2][1] * n.y + N[2][2] * n.z);
    };
    for (size_t i = 0; i + 2 < idx.size(); i += 3) {
        const uint32_t a = idx[i], b = idx[i + 1], c = idx[i + 2];
        if (a >= pos.size() || b >= pos.size() || c >= pos.size()) continue;
        BteTriangle t;
        std::memset(&t, 0, sizeof t);
        t.v0 = xformPoint(world, pos[a].x, pos[a].y, pos[a].z);
        t.v1 = xformPoint(world, pos[b].x, pos[b].y, pos[b].z);
        t.v2 = xformPoint(world, pos[c].x, pos[c].y, pos[c].z);
        if (a < nrm.size() && b < nrm.size() && c < nrm.size()) {
            t.n0 = xn(nrm[a]);
            t.n1 = xn(nrm[b]);
            t.n2 = xn(nrm[c]);
        } else {
            const float e1x = t.v1.x - t.v0.x, e1y = t.v1.y - t.v0.y, e1z = t.v1.z - t.v0.z;
            const float e2x = t.v2.x - t.v0.x, e2y = t.v2.y - t.v0.y, e2z = t.v2.z - t.v0.z;
            BteFloat3 g = {e1y * e2z - e1z * e2y, e1z * e2x - e1x * e2z, e1x * e2y - e1y * e2x};
            const float len = std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
            if (len > 1e-20f) {
                g.x /= len;
                g.y /= len;
                g.z /= len;
            }
            t.n0 = t.n1 = t.n2 = g;
        }
        auto uv = [&](uint32_t k) { return k < uvs.size() ? uvs[k] : BteFloat2{0.0f, 0.0f}; };
        t.uv0 = uv(a);
        t.uv1 = uv(b);
        t.uv2 = uv(c);
        out.triangles.push_back(t);
        out.materialIds.push_back(slot);
    }
}

bool walk(const Doc& d, int ni, const M4& parent, const std::vector<int>& slotOf, int defaultSlot,
          const std::set<uint64_t>& skip, GltfModel& out, int depth) {
    const JValue* node = d.item("nodes", ni);
    if (!node) return true;
    if (depth > kMaxDepth) return false;
    if (skipNode(*node)) return true;
    const M4 world = mul(parent, nodeMatrix(*node));
    const JValue* mesh = d.item("meshes", node->integer("mesh", -1));
    const JValue* prims = mesh ? mesh->get("primitives") : nullptr;
    if (prims && prims->isArray()) {
        for (size_t pi = 0; pi < prims->arr.size(); ++pi) {
            const JValue& prim = prims->arr[pi];
            if (skipPrim(d, prim) || skip.count(primKey(ni, (int)pi))) continue;
            const int m = prim.integer("material", -1);
            const int slot = (m >= 0 && (size_t)m < slotOf.size()) ? slotOf[(size_t)m] : defaultSlot;
            emitPrimitive(d, prim, world, slot, out);
        }
    }
    const JValue* ch = node->get("children");
    if (ch && ch->isArray())
        for (const JValue& c : ch->arr)
            if (c.isNumber() && !walk(d, toIndex(c.num), world, slotOf, defaultSlot, skip, out, depth + 1)) return false;
    return true;
}

double extNumber(const JValue* ext, const char* key, double def) {
    return (ext && ext->isObject()) ? ext->number(key, def) : def;
}

GltfImage textureImage(const Doc& d, const std::vector<GltfImage>& images, const JValue* texInfo) {
    if (!texInfo) return {};
    const JValue* tex = d.item("textures", texInfo->integer("index", -1));
    if (!tex) return {};
    const int src = tex->integer("source", -1);
    if (src < 0 || (size_t)src >= images.size()) return {};
    return images[(size_t)src];
}

GltfSlot disney(const Doc& d, const JValue& mat, const std::vector<GltfImage>& images) {
    GltfSlot slot;
    const JValue* pbr = mat.get("pbrMetallicRoughness");
    float base[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float metallic = 1.0f, roughness = 1.0f;
    if (pbr) {
        const std::vector<double> bc = pbr->numbers("baseColorFactor");
        if (bc.size() >= 3)
            for (size_t k = 0; k < 4; ++k) base[k] = k < bc.size() ? (float)bc[k] : 1.0f;
        metallic = (float)pbr->number("metallicFactor", 1.0);
        roughness = (float)pbr->number("roughnessFactor", 1.0);
        slot.albedo = textureImage(d, images, pbr->get("baseColorTexture"));
        slot.metallicRoughness = textureImage(d, images, pbr->get("metallicRoughnessTexture"));
        if (const JValue* bt = pbr->get("baseColorTexture"))
            if (bt->integer("texCoord", 0) != 0)
                std::fprintf(stderr, "[gltf] TEXCOORD_%d requested; using TEXCOORD_0\n", bt->integer("texCoord", 0));
    }
    // the base colour factor is baked into the albedo texture
    if (!slot.albedo.px.empty()) {
        std::vector<uint8_t>& p = slot.albedo.px;
        for (size_t i = 0; i < p.size(); ++i) {
            const float v = (p[i] / 255.0f) * base[i & 3] * 255.0f + 0.5f;
            p[i] = (uint8_t)std::min(std::max(v, 0.0f), 255.0f);
        }
    }
    const JValue* ext = mat.get("extensions");
    auto extension = [&](const char* name) { return ext ? ext->get(name) : nullptr; };
    float specTrans = 0.0f, sheen = 0.0f, sheenTint = 0.5f, clearcoat = 0.0f, clearcoatGloss = 1.0f, eta = 1.5f;
    if (const JValue* e = extension("KHR_materials_ior")) eta = (float)extNumber(e, "ior", eta);
    if (const JValue* e = extension("KHR_materials_transmission"))
        specTrans = (float)extNumber(e, "transmissionFactor", 0.0);
    if (const JValue* e = extension("KHR_materials_clearcoat")) {
        clearcoat = (float)extNumber(e, "clearcoatFactor", 0.0);
        clearcoatGloss = 1.0f - (float)extNumber(e, "clearcoatRoughnessFactor", 0.0);
    }
    if (const JValue* e = extension("KHR_materials_sheen")) {
        const JValue* sc = e->isObject() ? e->get("sheenColorFactor") : nullptr;
        if (sc && sc->isArray() && sc->arr.size() >= 3) {
            auto num = [&](size_t k) { return sc->arr[k].isNumber() ? (float)sc->arr[k].num : 0.0f; };
            sheen = 0.2126f * num(0) + 0.7152f * num(1) + 0.0722f * num(2);
            sheenTint = 1.0f;
        }
    }
    float em[3] = {0.0f, 0.0f, 0.0f};
    const std::vector<double> ef = mat.numbers("emissiveFactor");
    if (ef.size() >= 3)
        for (int k = 0; k < 3; ++k) em[k] = (float)ef[(size_t)k];
    if (const JValue* e = extension("KHR_materials_emissive_strength")) {
        const float s = (float)extNumber(e, "emissiveStrength", 1.0);
        for (float& v : em) v *= s;
    }
    slot.emission = {em[0], em[1], em[2]};
    slot.isEmitter = (em[0] + em[1] + em[2]) > 1e-4f;
    const bool unlit = extension("KHR_materials_unlit") != nullptr;
    const bool textured = !slot.albedo.px.empty();
    BteBsdf& b = slot.bsdf;
    std::memset(&b, 0, sizeof b);
    b.type = BTE_BSDF_DISNEY;
    b.p0 = {textured ? 1.0f : base[0], textured ? 1.0f : base[1], textured ? 1.0f : base[2], unlit ? 1.0f : roughness};
    b.p1 = {unlit ? 0.0f : metallic, 0.5f, specTrans, 0.0f};
    b.p2 = {sheen, sheenTint, 0.0f, 0.0f};
    b.p3 = {clearcoat, clearcoatGloss, eta, 0.0f};
    return slot;
}

GltfSlot defaultSlot() {
    GltfSlot s;
    std::memset(&s.bsdf, 0, sizeof s.bsdf);
    s.bsdf.type = BTE_BSDF_DISNEY;
    s.bsdf.p0 = {0.8f, 0.8f, 0.8f, 0.5f};
    s.bsdf.p1 = {0.0f, 0.5f, 0.0f, 0.0f};
    s.bsdf.p2 = {0.0f, 0.5f, 0.0f, 0.0f};
    s.bsdf.p3 = {0.0f, 1.0f, 1.5f, 0.0f};
    return s;
}

uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

bool readDocument(const std::string& path, Doc& d) {
    std::vector<uint8_t> file;
    if (!readAll(path, file) || file.empty()) {
        d.err = "cannot read \"" + path + "\"";
        return false;
    }
    const size_t slash = path.find_last_of('/');
    d.dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    std::string json;
    std::vector<uint8_t> bin;
    bool haveBin = false;
    if (file.size() >= 12 && le32(file.data()) == 0x46546C67u) {   // "glTF" binary container
        if (le32(file.data() + 4) != 2) {
            d.err = "GLB version is not 2";
            return false;
        }
        const size_t total = std::min<size_t>(le32(file.data() + 8), file.size());
        size_t pos = 12;
        while (pos + 8 <= total) {
            const size_t len = le32(file.data() + pos);
            const uint32_t type = le32(file.data() + pos + 4);
            if (len > total - pos - 8) {
                d.err = "truncated GLB chunk";
                return false;
            }
            const uint8_t* body = file.data() + pos + 8;
            if (type == 0x4E4F534Au && json.empty()) json.assign((const char*)body, len);
            else if (type == 0x004E4942u && !haveBin) {
                bin.assign(body, body + len);
                haveBin = true;
            }
            pos += 8 + ((len + 3) & ~(size_t)3);
        }
        if (json.empty()) {
            d.err = "GLB without a JSON chunk";
            return false;
        }
    } else {
        json.assign((const char*)file.data(), file.size());
    }
    std::string jerr;
    if (!parse_json(json.data(), json.size(), d.j, jerr) || !d.j.isObject()) {
        d.err = jerr.empty() ? "the glTF JSON is not an object" : jerr;
        return false;
    }
    const JValue* asset = d.j.get("asset");
    const std::string version = asset ? asset->string("version") : std::string();
    if (!version.empty() && version[0] != '2') {
        d.err = "not a glTF 2.0 asset (version " + version + ")";
        return false;
    }
    for (size_t i = 0; i < d.count("buffers"); ++i) {
        const JValue& b = *d.item("buffers", (int)i);
        std::vector<uint8_t> data;
        const JValue* uri = b.get("uri");
        if (uri && uri->type == JValue::String) {
            if (!d.loadUri(uri->str, data)) return false;
        } else if (i == 0 && haveBin) {
            data = bin;
        } else {
            d.err = "buffer " + std::to_string(i) + " has no data";
            return false;
        }
        const double want = b.number("byteLength", 0.0);
        if ((double)data.size() < want) {
            d.err = "buffer " + std::to_string(i) + " is shorter than its byteLength";
            return false;
        }
        d.buffers.push_back(std::move(data));
    }
    return true;
}

}  // namespace

bool load_gltf(const std::string& path, const ModelPlacement& place, GltfModel& out, std::string& err) {
    out = GltfModel();
    Doc d;
    if (!readDocument(path, d)) {
        err = d.err;
        return false;
    }
    // every image is decoded up front (a failing one fails the load)
    std::vector<GltfImage> images(d.count("images"));
    for (size_t i = 0; i < images.size(); ++i) {
        const JValue& im = *d.item("images", (int)i);
        std::vector<uint8_t> bytes;
        const JValue* uri = im.get("uri");
        if (uri && uri->type == JValue::String) {
            if (!d.loadUri(uri->str, bytes)) {
                err = "image " + std::to_string(i) + ": " + d.err;
                return false;
            }
        } else if (const JValue* bv = d.item("bufferViews", im.integer("bufferView", -1))) {
            const int b = bv->integer("buffer", -1);
            size_t off = 0, len = 0;
            if (b < 0 || (size_t)b >= d.buffers.size() || !toSize(bv->number("byteOffset", 0.0), off) ||
                !toSize(bv->number("byteLength", 0.0), len) || off > d.buffers[(size_t)b].size() ||
                len > d.buffers[(size_t)b].size() - off) {
                err = "image " + std::to_string(i) + ": buffer view out of range";
                return false;
            }
            bytes.assign(d.buffers[(size_t)b].begin() + (std::ptrdiff_t)off,
                         d.buffers[(size_t)b].begin() + (std::ptrdiff_t)(off + len));
        } else {
            err = "image " + std::to_string(i) + " has neither a uri nor a bufferView";
            return false;
        }
        std::string ierr;
        if (!decode_image(bytes.data(), bytes.size(), images[i].width, images[i].height, images[i].px, ierr)) {
            err = "image " + std::to_string(i) + ": " + ierr;
            return false;
        }
    }
    std::vector<int> slotOf(d.count("materials"), 0);
    for (size_t i = 0; i < slotOf.size(); ++i) {
        slotOf[i] = (int)out.materials.size();
        out.materials.push_back(disney(d, *d.item("materials", (int)i), images));
    }
    const int defSlot = (int)out.materials.size();
    out.materials.push_back(defaultSlot());

    const float d2r = 3.14159265358979323846f / 180.0f;
    M4 placeM = translation(place.pos[0], place.pos[1], place.pos[2]);
    placeM = mul(placeM, rotation(2, place.rotDeg[2] * d2r));
    placeM = mul(placeM, rotation(1, place.rotDeg[1] * d2r));
    placeM = mul(placeM, rotation(0, place.rotDeg[0] * d2r));
    placeM = mul(placeM, scaling(place.scale[0], place.scale[1], place.scale[2]));

    std::vector<int> roots;
    const int sceneIndex = d.j.integer("scene", -1) >= 0 ? d.j.integer("scene", -1) : 0;
    if (const JValue* sc = d.item("scenes", sceneIndex)) {
        for (double n : sc->numbers("nodes")) roots.push_back(toIndex(n));
    } else {
        for (size_t i = 0; i < d.count("nodes"); ++i) roots.push_back((int)i);
    }
    std::vector<PrimStat> stats;
    for (int r : roots)
        if (!gatherStats(d, r, placeM, stats, 0)) {
            err = "node hierarchy deeper than 1024 levels (a cycle?)";
            return false;
        }
    const std::set<uint64_t> skip = outliers(stats);
    if (!skip.empty())
        std::fprintf(stderr, "[gltf] skipping %zu detached/outlier mesh(es) in \"%s\"\n", skip.size(), path.c_str());
    for (int r : roots)
        if (!walk(d, r, placeM, slotOf, defSlot, skip, out, 0)) {
            err = "node hierarchy deeper than 1024 levels (a cycle?)";
            return false;
        }
    if (!out.materialIds.empty() &&
        std::find(out.materialIds.begin(), out.materialIds.end(), defSlot) == out.materialIds.end() &&
        defSlot == (int)out.materials.size() - 1)
        out.materials.pop_back();
    if (out.triangles.empty()) {
        err = "no triangles in \"" + path + "\"";
        return false;
    }
    return true;
}

}  // namespace bte
