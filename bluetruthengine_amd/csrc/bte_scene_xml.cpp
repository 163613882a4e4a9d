// bte_scene_xml.cpp — the reference's scene host layer: scene.xml parsing,
// OBJ import, BSDF packing, emitter tables, camera basis, spotlights and the
// Radiance .hdr environment, producing the flattened arrays the renderer's
// init entry points take.
//
// Follows, step for step where the result depends on it:
//   loadSceneDescription            src/scene.cpp:83-523 (tag/attribute scan,
//                                   defaults include/scene.h:9-147)
//   loadTrianglesFromObj            src/main.cpp:267-401 (normal / planar-UV
//                                   fallbacks), transforms include/mesh_transform.h:9-80
//   camera basis                    src/main.cpp:534-561
//   BSDF packing                    src/main.cpp:630-665
//   triangle / emitter tables       src/main.cpp:669-758 (triangleAreaHost :469-483)
//   spotlights                      src/main.cpp:879-904
//   environment map                 src/main.cpp:907-930 (stbi_loadf of an RGBE file)
//   albedo textures                 src/main.cpp:407-474, 600-612, 822-837
//                                   (loadTextureAverage, loadImageRGBA, one
//                                   image per material, no metallic-roughness
//                                   maps for XML materials)
//   glTF / GLB <model> entries      src/scene.cpp:274-316, src/main.cpp:760-818
//                                   (loader in bte_gltf.cpp)
// Differences by design: errors return a status and message instead of
// exit(); relative paths resolve against a caller-given directory.  Textures
// are decoded by the
// PNG and baseline JPEG decoders in bte_png.cpp / bte_image.cpp: a texture
// path that does not exist falls back to the BSDF colour (white average) as
// in the reference, while a file that exists but cannot be decoded (another
// format, progressive JPEG, corrupt data) is an error rather than a silently
// untextured material.  Numbers are
// parsed with strtod and rounded to float (OBJ: the vendored tinyobjloader's
// fast_float path rounds the same way; XML: std::stof).
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bte_api.h"
#include "bte_gltf.h"
#include "bte_image.h"

extern "C" {
struct BteScene {
    std::vector<BteTriangle> tris;
    std::vector<BteFloat3> materials;
    std::vector<int> triMaterialIds, triBsdfIds;
    std::vector<BteBsdf> bsdfs;
    std::vector<BteFloat3> triEmission;
    std::vector<uint8_t> triEmitterFlags;
    std::vector<BteEmitter> emitters;
    std::vector<int> emTriIdx;
    std::vector<float> emTriCdf, sceneCdf;
    std::vector<BteSpotlight> spots;
    std::vector<float> env;   // RGBA float, row 0 = top
    int envW = 0, envH = 0;
    float envIntensity = 0.0f, envRotation = 0.0f;
    BteCamera camera{};
    int width = 0, height = 0;
    // one albedo and one metallic-roughness image per material (RGBA8, row 0
    // = top); empty = no texture
    std::vector<std::vector<uint8_t>> albedo, mr;
    std::vector<const uint8_t*> albedoPtrs, mrPtrs;
    std::vector<int> albedoW, albedoH, mrW, mrH;
};
}

namespace {

char g_sceneErr[1024] = "";

int sceneFail(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int sceneFail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_sceneErr, sizeof(g_sceneErr), fmt, ap);
    va_end(ap);
    return BTE_ERR_INVALID;
}

bool readFile(const std::string& path, std::string& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    const bool ok = n <= 0 || std::fread(&out[0], 1, (size_t)n, f) == (size_t)n;
    std::fclose(f);
    return ok;
}

std::string joinPath(const std::string& base, const std::string& rel) {
    if (rel.empty() || rel[0] == '/' || base.empty()) return rel;
    return base.back() == '/' ? base + rel : base + "/" + rel;
}

bool fileExists(const std::string& p) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// ---- the reference's lenient tag scan ---------------------------------------
// The text of the next "<tag ...>" (up to, not including, '>') from `from`;
// comments are not skipped (the reference's scan does not skip them either).
bool nextTag(const std::string& xml, const char* tag, size_t& from, std::string& content) {
    const std::string open = std::string("<") + tag;
    const size_t p = xml.find(open, from);
    if (p == std::string::npos) return false;
    const size_t e = xml.find('>', p);
    if (e == std::string::npos) return false;
    content = xml.substr(p, e - p);
    from = e + 1;
    return true;
}

// Value of attr="..." anywhere in the tag text (first occurrence), or empty.
std::string attr(const std::string& tagText, const char* name) {
    const std::string key = std::string(name) + "=\"";
    const size_t p = tagText.find(key);
    if (p == std::string::npos) return {};
    const size_t b = p + key.size();
    const size_t e = tagText.find('"', b);
    if (e == std::string::npos) return {};
    return tagText.substr(b, e - b);
}

// std::stof semantics: leading whitespace, longest valid prefix, error when
// nothing converts or the value is out of float range.
bool toFloat(const std::string& s, float& out) {
    const char* c = s.c_str();
    char* end = nullptr;
    errno = 0;
    const float v = std::strtof(c, &end);
    if (end == c || errno == ERANGE) return false;
    out = v;
    return true;
}

bool toInt(const std::string& s, int& out) {
    const char* c = s.c_str();
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(c, &end, 10);
    if (end == c || errno == ERANGE || v < -2147483647L - 1 || v > 2147483647L) return false;
    out = (int)v;
    return true;
}

struct Transform {
    float posX = 0, posY = 0, posZ = 0, rotX = 0, rotY = 0, rotZ = 0, scaleX = 1, scaleY = 1, scaleZ = 1;
};

// Scale -> Rx -> Ry -> Rz -> translate (degrees), include/mesh_transform.h:9-46
BteFloat3 transformPoint(BteFloat3 p, const Transform& t) {
    float x = p.x * t.scaleX, y = p.y * t.scaleY, z = p.z * t.scaleZ;
    const float d2r = 3.14159265358979323846f / 180.0f;
    const float rx = t.rotX * d2r, ry = t.rotY * d2r, rz = t.rotZ * d2r;
    {
        const float c = std::cos(rx), s = std::sin(rx);
        const float y2 = y * c - z * s, z2 = y * s + z * c;
        y = y2; z = z2;
    }
    {
        const float c = std::cos(ry), s = std::sin(ry);
        const float x2 = x * c + z * s, z2 = -x * s + z * c;
        x = x2; z = z2;
    }
    {
        const float c = std::cos(rz), s = std::sin(rz);
        const float x2 = x * c - y * s, y2 = x * s + y * c;
        x = x2; y = y2;
    }
    return {x + t.posX, y + t.posY, z + t.posZ};
}

// rotations only, then renormalised (include/mesh_transform.h:48-80)
BteFloat3 transformNormal(BteFloat3 n, const Transform& t) {
    float x = n.x, y = n.y, z = n.z;
    const float d2r = 3.14159265358979323846f / 180.0f;
    const float rx = t.rotX * d2r, ry = t.rotY * d2r, rz = t.rotZ * d2r;
    {
        const float c = std::cos(rx), s = std::sin(rx);
        const float y2 = y * c - z * s, z2 = y * s + z * c;
        y = y2; z = z2;
    }
    {
        const float c = std::cos(ry), s = std::sin(ry);
        const float x2 = x * c + z * s, z2 = -x * s + z * c;
        x = x2; z = z2;
    }
    {
        const float c = std::cos(rz), s = std::sin(rz);
        const float x2 = x * c - y * s, y2 = x * s + y * c;
        x = x2; y = y2;
    }
    const float len = std::sqrt(x * x + y * y + z * z);
    if (len > 1e-8f) { x /= len; y /= len; z /= len; }
    return {x, y, z};
}

// ---- OBJ --------------------------------------------------------------------
struct ObjIndex {
    int v, vt, vn;   // 0-based, -1 when absent
};

// One face-vertex token "v", "v/vt", "v//vn" or "v/vt/vn"; negative indices
// count back from the current end of the list.
bool parseFaceVertex(const char* tok, size_t nv, size_t nvt, size_t nvn, ObjIndex& out) {
    out = {-1, -1, -1};
    long vals[3] = {0, 0, 0};
    bool have[3] = {false, false, false};
    const char* p = tok;
    for (int k = 0; k < 3; ++k) {
        if (*p && *p != '/') {
            char* e = nullptr;
            vals[k] = std::strtol(p, &e, 10);
            if (e == p) return false;
            have[k] = true;
            p = e;
        }
        if (*p != '/') break;
        ++p;
    }
    const size_t counts[3] = {nv, nvt, nvn};
    int* dst[3] = {&out.v, &out.vt, &out.vn};
    for (int k = 0; k < 3; ++k) {
        if (!have[k]) continue;
        long i = vals[k];
        if (i > 0) i -= 1;
        else if (i < 0) i = (long)counts[k] + i;
        else return false;   // index 0 is invalid in OBJ
        if (i < 0) return false;
        *dst[k] = (int)i;
    }
    return have[0];
}

struct ObjData {
    std::vector<float> v, vn, vt;
    std::vector<ObjIndex> tri;   // 3 per triangle (polygons fan-triangulated)
};

int loadObj(const std::string& path, ObjData& o) {
    std::string text;
    if (!readFile(path, text)) return sceneFail("cannot open OBJ \"%s\"", path.c_str());
    size_t pos = 0;
    std::vector<ObjIndex> poly;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        size_t s = line.find_first_not_of(" \t");
        if (s == std::string::npos || line[s] == '#') continue;
        const char* c = line.c_str() + s;
        auto readFloats = [&](const char* q, int want, std::vector<float>& dst, float dflt) {
            for (int k = 0; k < want; ++k) {
                while (*q == ' ' || *q == '\t') ++q;
                char* e = nullptr;
                const double d = std::strtod(q, &e);
                if (e == q) { dst.push_back(dflt); continue; }
                dst.push_back((float)d);
                q = e;
            }
        };
        if (c[0] == 'v' && (c[1] == ' ' || c[1] == '\t')) {
            readFloats(c + 2, 3, o.v, 0.0f);
        } else if (c[0] == 'v' && c[1] == 'n' && (c[2] == ' ' || c[2] == '\t')) {
            readFloats(c + 3, 3, o.vn, 0.0f);
        } else if (c[0] == 'v' && c[1] == 't' && (c[2] == ' ' || c[2] == '\t')) {
            readFloats(c + 3, 2, o.vt, 0.0f);
        } else if (c[0] == 'f' && (c[1] == ' ' || c[1] == '\t')) {
            poly.clear();
            const char* q = c + 2;
            while (*q) {
                while (*q == ' ' || *q == '\t') ++q;
                if (!*q) break;
                const char* b = q;
                while (*q && *q != ' ' && *q != '\t') ++q;
                const std::string tok(b, q);
                ObjIndex ix;
                if (!parseFaceVertex(tok.c_str(), o.v.size() / 3, o.vt.size() / 2, o.vn.size() / 3, ix))
                    return sceneFail("OBJ \"%s\": bad face vertex \"%s\"", path.c_str(), tok.c_str());
                if ((size_t)ix.v >= o.v.size() / 3)
                    return sceneFail("OBJ \"%s\": vertex index out of range", path.c_str());
                if (ix.vt >= 0 && (size_t)ix.vt >= o.vt.size() / 2) ix.vt = -1;
                if (ix.vn >= 0 && (size_t)ix.vn >= o.vn.size() / 3) ix.vn = -1;
                poly.push_back(ix);
            }
            for (size_t k = 1; k + 1 < poly.size(); ++k) {
                o.tri.push_back(poly[0]);
                o.tri.push_back(poly[k]);
                o.tri.push_back(poly[k + 1]);
            }
        }
    }
    if (o.tri.empty()) return sceneFail("no triangles found in OBJ \"%s\"", path.c_str());
    return BTE_OK;
}

// OBJ -> TriangleData (src/main.cpp:267-401)
int objTriangles(const std::string& path, const Transform& t, std::vector<BteTriangle>& out) {
    ObjData o;
    const int rc = loadObj(path, o);
    if (rc) return rc;
    const bool realUVs = o.vt.size() > 2;   // more than one texcoord
    auto planarUV = [](BteFloat3 p, BteFloat3 n) -> BteFloat2 {
        const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
        if (ay >= ax && ay >= az) return {p.x, p.z};
        if (ax >= ay && ax >= az) return {p.y, p.z};
        return {p.x, p.y};
    };
    for (size_t i = 0; i + 2 < o.tri.size(); i += 3) {
        const ObjIndex a = o.tri[i], b = o.tri[i + 1], c = o.tri[i + 2];
        BteTriangle tr;
        std::memset(&tr, 0, sizeof tr);
        auto V = [&](int k) { return BteFloat3{o.v[3 * (size_t)k], o.v[3 * (size_t)k + 1], o.v[3 * (size_t)k + 2]}; };
        auto N = [&](int k) { return BteFloat3{o.vn[3 * (size_t)k], o.vn[3 * (size_t)k + 1], o.vn[3 * (size_t)k + 2]}; };
        tr.v0 = transformPoint(V(a.v), t);
        tr.v1 = transformPoint(V(b.v), t);
        tr.v2 = transformPoint(V(c.v), t);
        const bool hasN = a.vn >= 0 && b.vn >= 0 && c.vn >= 0 && !o.vn.empty();
        if (hasN) {
            tr.n0 = transformNormal(N(a.vn), t);
            tr.n1 = transformNormal(N(b.vn), t);
            tr.n2 = transformNormal(N(c.vn), t);
        } else {
            const BteFloat3 e1 = {tr.v1.x - tr.v0.x, tr.v1.y - tr.v0.y, tr.v1.z - tr.v0.z};
            const BteFloat3 e2 = {tr.v2.x - tr.v0.x, tr.v2.y - tr.v0.y, tr.v2.z - tr.v0.z};
            BteFloat3 g = {e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x};
            const float len = std::sqrt(g.x * g.x + g.y * g.y + g.z * g.z);
            if (len > 1e-8f) { g.x /= len; g.y /= len; g.z /= len; }
            tr.n0 = tr.n1 = tr.n2 = g;
        }
        if (realUVs) {
            auto T = [&](int k) {
                return k < 0 ? BteFloat2{0.0f, 0.0f} : BteFloat2{o.vt[2 * (size_t)k], o.vt[2 * (size_t)k + 1]};
            };
            tr.uv0 = T(a.vt);
            tr.uv1 = T(b.vt);
            tr.uv2 = T(c.vt);
        } else {
            tr.uv0 = planarUV(tr.v0, tr.n0);
            tr.uv1 = planarUV(tr.v1, tr.n0);
            tr.uv2 = planarUV(tr.v2, tr.n0);
        }
        out.push_back(tr);
    }
    return BTE_OK;
}

float triangleArea(const BteTriangle& t) {   // src/main.cpp:469-483
    const float ax = t.v1.x - t.v0.x, ay = t.v1.y - t.v0.y, az = t.v1.z - t.v0.z;
    const float bx = t.v2.x - t.v0.x, by = t.v2.y - t.v0.y, bz = t.v2.z - t.v0.z;
    const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
}

// ---- Radiance RGBE (.hdr) -------------------------------------------------------
// Header lines up to an empty line ("#?RADIANCE"/"#?RGBE" first,
// FORMAT=32-bit_rle_rgbe), then "-Y H +X W"; scanlines flat or new-style
// run-length encoded (0x02 0x02 hi lo, then each of the 4 components as runs
// (count > 128: repeat next byte count-128 times) or literals).  Decoding as
// stb_image's stbi_loadf: c * 2^(e - 136) per channel, 0 when e = 0.
int loadHdr(const std::string& path, std::vector<float>& rgba, int& W, int& H) {
    std::string d;
    if (!readFile(path, d)) return sceneFail("cannot open environment map \"%s\"", path.c_str());
    size_t p = 0;
    auto line = [&](std::string& out) -> bool {
        const size_t e = d.find('\n', p);
        if (e == std::string::npos) return false;
        out = d.substr(p, e - p);
        p = e + 1;
        return true;
    };
    std::string l;
    if (!line(l) || (l != "#?RADIANCE" && l != "#?RGBE")) return sceneFail("\"%s\" is not a Radiance .hdr file", path.c_str());
    bool rgbe = false;
    for (;;) {
        if (!line(l)) return sceneFail("\"%s\": truncated header", path.c_str());
        if (l.empty()) break;
        if (l == "FORMAT=32-bit_rle_rgbe") rgbe = true;
    }
    if (!rgbe) return sceneFail("\"%s\": unsupported .hdr format (need 32-bit_rle_rgbe)", path.c_str());
    if (!line(l)) return sceneFail("\"%s\": missing resolution line", path.c_str());
    int h = 0, w = 0;
    char ys[3] = {0}, xs[3] = {0};
    if (std::sscanf(l.c_str(), "%2s %d %2s %d", ys, &h, xs, &w) != 4 || std::strcmp(ys, "-Y") != 0 ||
        std::strcmp(xs, "+X") != 0 || w <= 0 || h <= 0)
        return sceneFail("\"%s\": unsupported resolution line \"%s\"", path.c_str(), l.c_str());
    rgba.assign((size_t)w * h * 4, 0.0f);
    std::vector<unsigned char> scan((size_t)w * 4);
    const unsigned char* b = reinterpret_cast<const unsigned char*>(d.data());
    const size_t n = d.size();
    for (int y = 0; y < h; ++y) {
        const bool rle = w >= 8 && w < 32768 && p + 4 <= n && b[p] == 2 && b[p + 1] == 2 && !(b[p + 2] & 0x80);
        if (rle) {
            if (((int)b[p + 2] << 8 | b[p + 3]) != w) return sceneFail("\"%s\": bad scanline width", path.c_str());
            p += 4;
            for (int k = 0; k < 4; ++k) {
                int x = 0;
                while (x < w) {
                    if (p >= n) return sceneFail("\"%s\": truncated data", path.c_str());
                    int count = b[p++];
                    if (count > 128) {
                        count -= 128;
                        if (p >= n || x + count > w) return sceneFail("\"%s\": bad run", path.c_str());
                        const unsigned char v = b[p++];
                        for (int i = 0; i < count; ++i) scan[(size_t)(x++) * 4 + k] = v;
                    } else {
                        if (count == 0 || p + count > n || x + count > w)
                            return sceneFail("\"%s\": bad literal run", path.c_str());
                        for (int i = 0; i < count; ++i) scan[(size_t)(x++) * 4 + k] = b[p++];
                    }
                }
            }
        } else {
            if (p + (size_t)w * 4 > n) return sceneFail("\"%s\": truncated data", path.c_str());
            std::memcpy(scan.data(), b + p, (size_t)w * 4);
            p += (size_t)w * 4;
        }
        for (int x = 0; x < w; ++x) {
            const unsigned char* q = &scan[(size_t)x * 4];
            float* o = &rgba[((size_t)y * w + x) * 4];
            if (q[3]) {
                const float f1 = std::ldexp(1.0f, (int)q[3] - (128 + 8));
                o[0] = q[0] * f1;
                o[1] = q[1] * f1;
                o[2] = q[2] * f1;
            }
            o[3] = 1.0f;
        }
    }
    W = w;
    H = h;
    return BTE_OK;
}

struct BsdfDesc {
    std::string name, type;
    float albedo[3] = {1, 1, 1};
    float intIOR = 1.5f, extIOR = 1.0f, alpha = 0.1f;
    float base[3] = {0.5f, 0.5f, 0.5f};
    float roughness = 0.1f, metallic = 0, specular = 0.5f, specTrans = 0, specTint = 0, sheen = 0, sheenTint = 0.5f,
          subsurface = 0, anisotropic = 0, clearcoat = 0, clearcoatGloss = 1, eta = 1.5f;
};
struct MaterialDesc {
    std::string name, albedoTexture, bsdf;
};
struct MeshDesc {
    std::string filename, material;
    bool isEmitter = false;
    float radiance[3] = {0, 0, 0};
    Transform t;
};

int parse(const std::string& path, const std::string& baseDir, BteScene& sc) {
    std::string xml;
    if (!readFile(path, xml) || xml.empty()) return sceneFail("scene file \"%s\" is empty or unreadable", path.c_str());
    std::vector<BsdfDesc> bsdfs;
    std::vector<MaterialDesc> mats;
    std::vector<MeshDesc> meshes;
    std::string tag;
    // <bsdf> (src/scene.cpp:96-164)
    for (size_t from = 0; nextTag(xml, "bsdf", from, tag);) {
        BsdfDesc b;
        b.name = attr(tag, "name");
        b.type = attr(tag, "type");
        if (b.name.empty() || b.type.empty()) return sceneFail("<bsdf> missing required `name` and/or `type`");
        bool ok = true;
        auto rf = [&](const char* a, float& v) {
            const std::string s = attr(tag, a);
            if (!s.empty() && !toFloat(s, v)) ok = false;
        };
        if (b.type == "diffuse" || b.type == "mirror") {
            rf("albedoR", b.albedo[0]); rf("albedoG", b.albedo[1]); rf("albedoB", b.albedo[2]);
        } else if (b.type == "dielectric") {
            rf("intIOR", b.intIOR); rf("extIOR", b.extIOR);
        } else if (b.type == "microfacet") {
            rf("albedoR", b.albedo[0]); rf("albedoG", b.albedo[1]); rf("albedoB", b.albedo[2]);
            rf("alpha", b.alpha); rf("intIOR", b.intIOR); rf("extIOR", b.extIOR);
        } else if (b.type == "disney") {
            rf("base_colorR", b.base[0]); rf("base_colorG", b.base[1]); rf("base_colorB", b.base[2]);
            rf("roughness", b.roughness); rf("metallic", b.metallic); rf("specular", b.specular);
            rf("specular_transmission", b.specTrans); rf("specular_tint", b.specTint); rf("sheen", b.sheen);
            rf("sheen_tint", b.sheenTint); rf("subsurface", b.subsurface); rf("anisotropic", b.anisotropic);
            rf("clearcoat", b.clearcoat); rf("clearcoat_gloss", b.clearcoatGloss); rf("eta", b.eta);
        } else {
            return sceneFail("<bsdf name=\"%s\"> unsupported type \"%s\"", b.name.c_str(), b.type.c_str());
        }
        if (!ok) return sceneFail("<bsdf name=\"%s\"> invalid float attribute", b.name.c_str());
        bsdfs.push_back(b);
    }
    // <material> (src/scene.cpp:167-196)
    for (size_t from = 0; nextTag(xml, "material", from, tag);) {
        MaterialDesc m{attr(tag, "name"), attr(tag, "albedoTexture"), attr(tag, "bsdf")};
        if (m.name.empty()) return sceneFail("<material> missing required attribute `name`");
        if (m.bsdf.empty()) return sceneFail("<material> \"%s\" missing required attribute `bsdf`", m.name.c_str());
        mats.push_back(m);
    }
    // <mesh> (src/scene.cpp:199-276)
    for (size_t from = 0; nextTag(xml, "mesh", from, tag);) {
        MeshDesc m;
        m.filename = attr(tag, "filename");
        m.material = attr(tag, "material");
        bool ok = true;
        auto rf = [&](const char* a, float& v) {
            const std::string s = attr(tag, a);
            if (!s.empty() && !toFloat(s, v)) ok = false;
        };
        const std::string em = attr(tag, "isEmitter");
        if (!em.empty()) {
            if (em == "1" || em == "true" || em == "True") m.isEmitter = true;
            else if (em == "0" || em == "false" || em == "False") m.isEmitter = false;
            else return sceneFail("<mesh filename=\"%s\"> invalid bool attribute `isEmitter`", m.filename.c_str());
        }
        rf("radianceR", m.radiance[0]); rf("radianceG", m.radiance[1]); rf("radianceB", m.radiance[2]);
        if (m.isEmitter && m.radiance[0] == 0.0f && m.radiance[1] == 0.0f && m.radiance[2] == 0.0f)
            m.radiance[0] = m.radiance[1] = m.radiance[2] = 10.0f;
        rf("posX", m.t.posX); rf("posY", m.t.posY); rf("posZ", m.t.posZ);
        rf("rotX", m.t.rotX); rf("rotY", m.t.rotY); rf("rotZ", m.t.rotZ);
        rf("scaleX", m.t.scaleX); rf("scaleY", m.t.scaleY); rf("scaleZ", m.t.scaleZ);
        if (!ok) return sceneFail("<mesh filename=\"%s\"> invalid float attribute", m.filename.c_str());
        if (m.filename.empty()) return sceneFail("<mesh> missing required attribute `filename`");
        if (m.material.empty()) return sceneFail("<mesh filename=\"%s\"> missing required attribute `material`", m.filename.c_str());
        meshes.push_back(m);
    }
    // <model> (src/scene.cpp:274-316): glTF / GLB drop-ins
    struct ModelDesc {
        std::string filename;
        bte::ModelPlacement place;
    };
    std::vector<ModelDesc> models;
    for (size_t from = 0; nextTag(xml, "model", from, tag);) {
        ModelDesc m;
        m.filename = attr(tag, "filename");
        bool ok = true;
        auto rf = [&](const char* a, float& v) {
            const std::string t = attr(tag, a);
            if (!t.empty() && !toFloat(t, v)) ok = false;
        };
        rf("posX", m.place.pos[0]); rf("posY", m.place.pos[1]); rf("posZ", m.place.pos[2]);
        rf("rotX", m.place.rotDeg[0]); rf("rotY", m.place.rotDeg[1]); rf("rotZ", m.place.rotDeg[2]);
        rf("scaleX", m.place.scale[0]); rf("scaleY", m.place.scale[1]); rf("scaleZ", m.place.scale[2]);
        if (!ok) return sceneFail("<model filename=\"%s\"> invalid float attribute", m.filename.c_str());
        if (m.filename.empty()) return sceneFail("<model> missing required attribute `filename`");
        models.push_back(m);
    }
    auto matIndex = [&](const std::string& n) {
        for (size_t i = 0; i < mats.size(); ++i)
            if (mats[i].name == n) return (int)i;
        return -1;
    };
    auto bsdfIndex = [&](const std::string& n) {
        for (size_t i = 0; i < bsdfs.size(); ++i)
            if (bsdfs[i].name == n) return (int)i;
        return -1;
    };
    for (const MeshDesc& m : meshes)
        if (matIndex(m.material) < 0) return sceneFail("mesh material reference \"%s\" not found", m.material.c_str());
    for (const MaterialDesc& m : mats)
        if (bsdfIndex(m.bsdf) < 0) return sceneFail("material bsdf reference \"%s\" not found", m.bsdf.c_str());
    // <spotlight> (src/scene.cpp:356-389)
    struct Spot { float pos[3] = {0, 0, 0}, dir[3] = {0, -1, 0}, rad[3] = {1, 1, 1}, intensity = 1, inner = 0, outer = 180; };
    std::vector<Spot> spots;
    for (size_t from = 0; nextTag(xml, "spotlight", from, tag);) {
        Spot s;
        bool ok = true;
        auto rf = [&](const char* a, float& v) {
            const std::string t = attr(tag, a);
            if (!t.empty() && !toFloat(t, v)) ok = false;
        };
        rf("posX", s.pos[0]); rf("posY", s.pos[1]); rf("posZ", s.pos[2]);
        rf("dirX", s.dir[0]); rf("dirY", s.dir[1]); rf("dirZ", s.dir[2]);
        rf("radianceR", s.rad[0]); rf("radianceG", s.rad[1]); rf("radianceB", s.rad[2]);
        rf("intensity", s.intensity); rf("innerConeAngle", s.inner); rf("outerConeAngle", s.outer);
        if (!ok) return sceneFail("<spotlight> invalid float attribute");
        spots.push_back(s);
    }
    // <environment> (src/scene.cpp:392-420)
    std::string envPath;
    float envIntensity = 1.0f, envRotation = 0.0f;
    {
        size_t from = 0;
        if (nextTag(xml, "environment", from, tag)) {
            envPath = attr(tag, "map");
            if (envPath.empty()) return sceneFail("<environment> missing required `map` attribute");
            const std::string si = attr(tag, "intensity"), sr = attr(tag, "rotation");
            if (!si.empty() && !toFloat(si, envIntensity)) return sceneFail("<environment> invalid `intensity`");
            if (!sr.empty() && !toFloat(sr, envRotation)) return sceneFail("<environment> invalid `rotation`");
        }
    }
    // <window> (src/scene.cpp:423-433): first <window ...> tag
    {
        size_t from = 0;
        if (!nextTag(xml, "window", from, tag)) return sceneFail("<window> missing required `width` and/or `height`");
        const std::string w = attr(tag, "width"), h = attr(tag, "height");
        if (w.empty() || h.empty()) return sceneFail("<window> missing required `width` and/or `height`");
        sc.width = 1280;
        sc.height = 720;
        (void)toInt(w, sc.width);
        (void)toInt(h, sc.height);
    }
    // <camera> (src/scene.cpp:436-468), basis as src/main.cpp:534-561
    {
        size_t from = 0;
        if (!nextTag(xml, "camera", from, tag)) return sceneFail("no <camera ...> tag found");
        const char* names[10] = {"eyeX", "eyeY", "eyeZ", "lookAtX", "lookAtY", "lookAtZ", "upX", "upY", "upZ", "fovY"};
        float v[10];
        for (int i = 0; i < 10; ++i) {
            const std::string s = attr(tag, names[i]);
            if (s.empty()) return sceneFail("<camera> missing required attribute `%s`", names[i]);
            if (!toFloat(s, v[i])) return sceneFail("<camera> invalid float value for `%s`", names[i]);
        }
        if (sc.height <= 0) return sceneFail("invalid window height");
        auto norm = [](float x, float y, float z, float o[3]) {   // v * inversesqrt(dot(v, v))
            const float r = 1.0f / std::sqrt(x * x + y * y + z * z);
            o[0] = x * r; o[1] = y * r; o[2] = z * r;
        };
        float f[3], rt[3], up[3];
        norm(v[3] - v[0], v[4] - v[1], v[5] - v[2], f);
        const float cx = f[1] * v[8] - v[7] * f[2], cy = f[2] * v[6] - v[8] * f[0], cz = f[0] * v[7] - v[6] * f[1];
        norm(cx, cy, cz, rt);
        up[0] = rt[1] * f[2] - f[1] * rt[2];
        up[1] = rt[2] * f[0] - f[2] * rt[0];
        up[2] = rt[0] * f[1] - f[0] * rt[1];
        if (!(std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]) >= 1e-6f))
            return sceneFail("invalid camera basis: `up` is colinear with the view direction");
        sc.camera.origin = {v[0], v[1], v[2]};
        sc.camera.forward = {f[0], f[1], f[2]};
        sc.camera.right = {rt[0], rt[1], rt[2]};
        sc.camera.up = {up[0], up[1], up[2]};
        sc.camera.fovYRadians = v[9] * 0.01745329251994329576923690768489f;
        sc.camera.aspect = (float)sc.width / (float)sc.height;
    }
    // src/scene.cpp:498-512: materials and BSDFs are required by OBJ meshes only
    if (meshes.empty() && models.empty()) return sceneFail("no <mesh> or <model> elements found");
    if (!meshes.empty() && mats.empty()) return sceneFail("no <material ...> elements found");
    if (!meshes.empty() && bsdfs.empty()) return sceneFail("no <bsdf ...> elements found");

    // materials (src/main.cpp:600-612): the texture's average colour, which the
    // integrator never reads, and the RGBA8 image itself, which overrides the
    // BSDF colour at shading time; a missing file falls back to white / none
    for (const MaterialDesc& m : mats) {
        std::vector<uint8_t> px;
        int tw = 1, th = 1;
        BteFloat3 avg = {1.0f, 1.0f, 1.0f};
        const std::string tpath = joinPath(baseDir, m.albedoTexture);
        if (!m.albedoTexture.empty() && fileExists(tpath)) {
            std::string bytes, err;
            if (!readFile(tpath, bytes))
                return sceneFail("material \"%s\": cannot read albedo texture \"%s\"", m.name.c_str(),
                                 tpath.c_str());
            if (!bte::decode_image(reinterpret_cast<const uint8_t*>(bytes.data()), bytes.size(), tw, th, px, err))
                return sceneFail("material \"%s\": albedo texture \"%s\": %s", m.name.c_str(), tpath.c_str(),
                                 err.c_str());
            // loadTextureAverage (src/main.cpp:407-440): binary64 sums of byte/255
            double acc[3] = {0.0, 0.0, 0.0};
            const size_t n = (size_t)tw * (size_t)th;
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) acc[k] += px[4 * i + (size_t)k] / 255.0;
            avg = {(float)(acc[0] / (double)n), (float)(acc[1] / (double)n), (float)(acc[2] / (double)n)};
        }
        sc.materials.push_back(avg);
        sc.albedo.push_back(std::move(px));
        sc.albedoW.push_back(tw);
        sc.albedoH.push_back(th);
        sc.mr.emplace_back();   // XML materials have no metallic-roughness map
        sc.mrW.push_back(1);
        sc.mrH.push_back(1);
    }
    // BSDF table (src/main.cpp:630-665)
    for (const BsdfDesc& b : bsdfs) {
        BteBsdf d;
        std::memset(&d, 0, sizeof d);
        if (b.type == "diffuse") {
            d.type = BTE_BSDF_DIFFUSE;
            d.p0 = {b.albedo[0], b.albedo[1], b.albedo[2], 0.0f};
        } else if (b.type == "dielectric") {
            d.type = BTE_BSDF_DIELECTRIC;
            d.p1 = {b.intIOR, b.extIOR, 0.0f, 0.0f};
        } else if (b.type == "mirror") {
            d.type = BTE_BSDF_MIRROR;
            d.p0 = {b.albedo[0] > 0.0f ? b.albedo[0] : 1.0f, b.albedo[1] > 0.0f ? b.albedo[1] : 1.0f,
                    b.albedo[2] > 0.0f ? b.albedo[2] : 1.0f, 0.0f};
        } else if (b.type == "microfacet") {
            d.type = BTE_BSDF_MICROFACET;
            const float ks = 1.0f - std::fmax(b.albedo[0], std::fmax(b.albedo[1], b.albedo[2]));
            d.p0 = {b.albedo[0], b.albedo[1], b.albedo[2], 0.0f};
            d.p1 = {b.alpha, b.intIOR, b.extIOR, ks};
        } else {
            d.type = BTE_BSDF_DISNEY;
            d.p0 = {b.base[0], b.base[1], b.base[2], b.roughness};
            d.p1 = {b.metallic, b.specular, b.specTrans, b.specTint};
            d.p2 = {b.sheen, b.sheenTint, b.subsurface, b.anisotropic};
            d.p3 = {b.clearcoat, b.clearcoatGloss, b.eta, 0.0f};
        }
        sc.bsdfs.push_back(d);
    }
    // triangles and emitter tables (src/main.cpp:669-758)
    sc.sceneCdf.push_back(0.0f);
    float weightSum = 0.0f;
    for (const MeshDesc& m : meshes) {
        const int matId = matIndex(m.material);
        const int bsdfId = bsdfIndex(mats[(size_t)matId].bsdf);
        std::vector<BteTriangle> tris;
        const int rc = objTriangles(joinPath(baseDir, m.filename), m.t, tris);
        if (rc) return rc;
        const BteFloat3 emit = {m.radiance[0], m.radiance[1], m.radiance[2]};
        int triOffset = 0, cdfOffset = 0;
        float areaSum = 0.0f;
        if (m.isEmitter) {
            triOffset = (int)sc.emTriIdx.size();
            cdfOffset = (int)sc.emTriCdf.size();
            sc.emTriCdf.push_back(0.0f);
        }
        for (const BteTriangle& t : tris) {
            sc.tris.push_back(t);
            sc.triMaterialIds.push_back(matId);
            sc.triBsdfIds.push_back(bsdfId);
            sc.triEmitterFlags.push_back(m.isEmitter ? 1 : 0);
            if (m.isEmitter) {
                sc.triEmission.push_back(emit);
                sc.emTriIdx.push_back((int)sc.tris.size() - 1);
                areaSum += triangleArea(t);
                sc.emTriCdf.push_back(areaSum);
            } else {
                sc.triEmission.push_back({0.0f, 0.0f, 0.0f});
            }
        }
        if (m.isEmitter) {
            const float lum = 0.2126f * emit.x + 0.7152f * emit.y + 0.0722f * emit.z;
            const float weight = areaSum * lum;
            weightSum += weight;
            sc.sceneCdf.push_back(weightSum);
            BteEmitter e;
            std::memset(&e, 0, sizeof e);
            e.triIndexOffset = triOffset;
            e.triCount = (int)sc.emTriIdx.size() - triOffset;
            e.cdfOffset = cdfOffset;
            e.radiance = emit;
            e.areaSum = areaSum;
            e.powerWeight = weight;
            sc.emitters.push_back(e);
        }
    }
    // glTF / GLB models after every OBJ mesh (src/main.cpp:760-818): their
    // materials, BSDFs and maps are appended, and each emissive material is
    // one emitter over the model's triangles that use it
    for (const ModelDesc& md : models) {
        bte::GltfModel g;
        std::string err;
        const std::string mpath = joinPath(baseDir, md.filename);
        if (!bte::load_gltf(mpath, md.place, g, err))
            return sceneFail("<model filename=\"%s\">: %s", mpath.c_str(), err.c_str());
        const int matOffset = (int)sc.materials.size(), bsdfOffset = (int)sc.bsdfs.size();
        const int firstTri = (int)sc.tris.size();
        for (bte::GltfSlot& slot : g.materials) {
            sc.bsdfs.push_back(slot.bsdf);
            sc.materials.push_back({1.0f, 1.0f, 1.0f});
            sc.albedoW.push_back(slot.albedo.width > 0 ? slot.albedo.width : 1);
            sc.albedoH.push_back(slot.albedo.height > 0 ? slot.albedo.height : 1);
            sc.albedo.push_back(std::move(slot.albedo.px));
            sc.mrW.push_back(slot.metallicRoughness.width > 0 ? slot.metallicRoughness.width : 1);
            sc.mrH.push_back(slot.metallicRoughness.height > 0 ? slot.metallicRoughness.height : 1);
            sc.mr.push_back(std::move(slot.metallicRoughness.px));
        }
        for (size_t i = 0; i < g.triangles.size(); ++i) {
            const int local = g.materialIds[i];
            const bte::GltfSlot& slot = g.materials[(size_t)local];
            sc.tris.push_back(g.triangles[i]);
            sc.triMaterialIds.push_back(matOffset + local);
            sc.triBsdfIds.push_back(bsdfOffset + local);
            sc.triEmitterFlags.push_back(slot.isEmitter ? 1 : 0);
            sc.triEmission.push_back(slot.isEmitter ? slot.emission : BteFloat3{0.0f, 0.0f, 0.0f});
        }
        for (size_t mi = 0; mi < g.materials.size(); ++mi) {
            const bte::GltfSlot& slot = g.materials[mi];
            if (!slot.isEmitter) continue;
            const int triOffset = (int)sc.emTriIdx.size(), cdfOffset = (int)sc.emTriCdf.size();
            sc.emTriCdf.push_back(0.0f);
            float areaSum = 0.0f;
            const int matId = matOffset + (int)mi;
            for (int ti = firstTri; ti < (int)sc.tris.size(); ++ti) {
                if (sc.triMaterialIds[(size_t)ti] != matId) continue;
                sc.emTriIdx.push_back(ti);
                areaSum += triangleArea(sc.tris[(size_t)ti]);
                sc.emTriCdf.push_back(areaSum);
            }
            // a material with no area leaves its CDF start behind, as in the reference
            if (areaSum > 0.0f) {
                const BteFloat3 emit = slot.emission;
                const float lum = 0.2126f * emit.x + 0.7152f * emit.y + 0.0722f * emit.z;
                const float weight = areaSum * lum;
                weightSum += weight;
                sc.sceneCdf.push_back(weightSum);
                BteEmitter e;
                std::memset(&e, 0, sizeof e);
                e.triIndexOffset = triOffset;
                e.triCount = (int)sc.emTriIdx.size() - triOffset;
                e.cdfOffset = cdfOffset;
                e.radiance = emit;
                e.areaSum = areaSum;
                e.powerWeight = weight;
                sc.emitters.push_back(e);
            }
        }
    }
    if (sc.tris.empty()) return sceneFail("the scene has no triangles after loading meshes / models");
    for (const std::vector<uint8_t>& px : sc.albedo) sc.albedoPtrs.push_back(px.empty() ? nullptr : px.data());
    for (const std::vector<uint8_t>& px : sc.mr) sc.mrPtrs.push_back(px.empty() ? nullptr : px.data());
    // spotlights (src/main.cpp:879-904)
    for (const Spot& s : spots) {
        BteSpotlight sp;
        std::memset(&sp, 0, sizeof sp);
        sp.position = {s.pos[0], s.pos[1], s.pos[2]};
        float dx = s.dir[0], dy = s.dir[1], dz = s.dir[2];
        const float dl = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (dl > 1e-6f) { dx /= dl; dy /= dl; dz /= dl; }
        sp.direction = {dx, dy, dz};
        sp.radiance = {s.rad[0], s.rad[1], s.rad[2]};
        sp.intensity = s.intensity;
        const float kPi = 3.14159265358979323846f;
        sp.innerConeCosine = std::cos(s.inner * kPi / 180.0f);
        sp.outerConeCosine = std::cos(s.outer * kPi / 180.0f);
        sc.spots.push_back(sp);
    }
    // environment (src/main.cpp:907-930)
    if (!envPath.empty()) {
        const int rc = loadHdr(joinPath(baseDir, envPath), sc.env, sc.envW, sc.envH);
        if (rc) return rc;
        sc.envIntensity = envIntensity;
        sc.envRotation = envRotation;
    }
    return BTE_OK;
}

}  // namespace

extern "C" {

int bte_scene_load_xml(const char* path, const char* baseDir, BteScene** out) {
    if (!path || !out) return sceneFail("null argument");
    *out = nullptr;
    BteScene* sc = new BteScene();
    const int rc = parse(path, baseDir ? baseDir : "", *sc);
    if (rc) {
        delete sc;
        return rc;
    }
    *out = sc;
    return BTE_OK;
}

const char* bte_scene_last_error(void) { return g_sceneErr; }

int bte_scene_info(const BteScene* sc, BteSceneInfo* info) {
    if (!sc || !info) return sceneFail("null argument");
    info->width = sc->width;
    info->height = sc->height;
    info->triangleCount = (int)sc->tris.size();
    info->materialCount = (int)sc->materials.size();
    info->bsdfCount = (int)sc->bsdfs.size();
    info->emitterCount = (int)sc->emitters.size();
    info->emitterTriCount = (int)sc->emTriIdx.size();
    info->emitterCdfCount = (int)sc->emTriCdf.size();
    info->spotlightCount = (int)sc->spots.size();
    info->envWidth = sc->envW;
    info->envHeight = sc->envH;
    info->envIntensity = sc->envIntensity;
    info->envRotationDeg = sc->envRotation;
    info->camera = sc->camera;
    return BTE_OK;
}

int bte_scene_arrays(const BteScene* sc, BteSceneArrays* a) {
    if (!sc || !a) return sceneFail("null argument");
    a->triangles = sc->tris.data();
    a->materials = sc->materials.data();
    a->triangleMaterialIds = sc->triMaterialIds.data();
    a->triangleBsdfIds = sc->triBsdfIds.data();
    a->bsdfs = sc->bsdfs.data();
    a->triangleEmission = sc->triEmission.data();
    a->triangleEmitterFlags = sc->triEmitterFlags.data();
    a->emitters = sc->emitters.data();
    a->emitterTriIndices = sc->emTriIdx.data();
    a->emitterTriCdf = sc->emTriCdf.data();
    a->sceneEmitterCdf = sc->sceneCdf.data();
    a->spotlights = sc->spots.data();
    a->envRGBA = sc->env.empty() ? nullptr : sc->env.data();
    a->albedoTextures = sc->albedoPtrs.data();
    a->albedoTextureWidths = sc->albedoW.data();
    a->albedoTextureHeights = sc->albedoH.data();
    a->mrTextures = sc->mrPtrs.data();
    a->mrTextureWidths = sc->mrW.data();
    a->mrTextureHeights = sc->mrH.data();
    return BTE_OK;
}

// The reference's upload order (src/main.cpp:846-935).
int bte_scene_upload(const BteScene* sc) {
    if (!sc) return sceneFail("null scene");
    int rc;
    if ((rc = bte_init_scene(sc->tris.data(), (int)sc->tris.size(), sc->materials.data(), (int)sc->materials.size(),
                             sc->triMaterialIds.data(), sc->width, sc->height)))
        return rc;
    if ((rc = bte_init_camera(&sc->camera))) return rc;
    if ((rc = bte_init_triangle_emission(sc->triEmission.data(), (int)sc->triEmission.size()))) return rc;
    if ((rc = bte_init_triangle_emitter_flags(sc->triEmitterFlags.data(), (int)sc->triEmitterFlags.size()))) return rc;
    if (!sc->emitters.empty())
        rc = bte_init_emitter_table(sc->emitters.data(), (int)sc->emitters.size(), sc->emTriIdx.data(),
                                    (int)sc->emTriIdx.size(), sc->emTriCdf.data(), (int)sc->emTriCdf.size(),
                                    sc->sceneCdf.data());
    else
        rc = bte_init_emitter_table(nullptr, 0, nullptr, 0, nullptr, 0, nullptr);
    if (rc) return rc;
    if ((rc = bte_init_bsdfs(sc->bsdfs.data(), (int)sc->bsdfs.size(), sc->triBsdfIds.data(),
                             (int)sc->triBsdfIds.size())))
        return rc;
    if ((rc = bte_init_spotlights(sc->spots.empty() ? nullptr : sc->spots.data(), (int)sc->spots.size()))) return rc;
    // src/main.cpp:822-837: one albedo entry per material (null = none), no
    // metallic-roughness maps
    auto any = [](const std::vector<const uint8_t*>& v) {
        for (const uint8_t* p : v)
            if (p) return true;
        return false;
    };
    // the reference uploads both per-material tables (null entries included);
    // metallic-roughness maps are applied through the albedo table's count
    if (any(sc->albedoPtrs) || any(sc->mrPtrs))
        rc = bte_init_textures(sc->albedoPtrs.data(), sc->albedoW.data(), sc->albedoH.data(),
                               (int)sc->albedoPtrs.size());
    else
        rc = bte_init_textures(nullptr, nullptr, nullptr, 0);
    if (rc) return rc;
    if (any(sc->mrPtrs))
        rc = bte_init_mr_textures(sc->mrPtrs.data(), sc->mrW.data(), sc->mrH.data(), (int)sc->mrPtrs.size());
    else
        rc = bte_init_mr_textures(nullptr, nullptr, nullptr, 0);
    if (rc) return rc;
    if (sc->env.empty()) return bte_init_env_map(nullptr, 0, 0, 0.0f, 0.0f);
    return bte_init_env_map(sc->env.data(), sc->envW, sc->envH, sc->envIntensity, sc->envRotation);
}

void bte_scene_free(BteScene* sc) { delete sc; }

}  // extern "C"
