// bte_oracle.cpp — ORACLE (test infrastructure only; never linked into the
// product).  CPU restatement of the reference integrator; see bte_oracle.h for
// the list of reference files followed and the deliberate restatement choices.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).
// Every arithmetic expression keeps the reference's evaluation order, because
// the HIP kernels are required to reproduce these results bit-for-bit.

#include "bte_oracle.h"
#include "ora_math.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace ora {

using F2 = OraFloat2;
using F3 = OraFloat3;

// ---------------------------------------------------------------------------
// Vector helpers — include/rt_cuda_math.cuh:7-123
// ---------------------------------------------------------------------------
static inline F3 sub3(const F3& a, const F3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline F3 add3(const F3& a, const F3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline F3 mul3(const F3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline F3 mul3(const F3& a, const F3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline F3 div3(const F3& a, float s) {                       // rt_cuda_math.cuh:27-32
    if (s == 0.0f) return {0.0f, 0.0f, 0.0f};
    const float inv = 1.0f / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
static inline float dot3(const F3& a, const F3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline F3 cross3(const F3& a, const F3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline F3 normalize3(const F3& v) {                          // rt_cuda_math.cuh:49-55 (1/sqrt)
    const float len2 = dot3(v, v);
    if (len2 <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const float invLen = 1.0f / std::sqrt(len2);
    return mul3(v, invLen);
}
static inline float length3(const F3& v) { return std::sqrt(std::fmax(0.0f, dot3(v, v))); }
static inline float distance3(const F3& a, const F3& b) { return length3(sub3(a, b)); }
static inline float maxCoeff3(const F3& v) { return std::fmax(v.x, std::fmax(v.y, v.z)); }

struct Frame3 { F3 t, b, n; };
static inline Frame3 makeFrameFromNormal(const F3& normal) {        // rt_cuda_math.cuh:82-95
    Frame3 f;
    f.n = normalize3(normal);
    const F3 a = (std::fabs(f.n.z) < 0.999f) ? F3{0.0f, 0.0f, 1.0f} : F3{1.0f, 0.0f, 0.0f};
    f.t = normalize3(cross3(a, f.n));
    f.b = cross3(f.n, f.t);
    return f;
}
static inline F3 toLocal(const Frame3& f, const F3& v) { return {dot3(v, f.t), dot3(v, f.b), dot3(v, f.n)}; }
static inline F3 toWorld(const Frame3& f, const F3& v) {
    return add3(add3(mul3(f.t, v.x), mul3(f.b, v.y)), mul3(f.n, v.z));
}
static inline F3 toLocalFromNormal(const F3& n, const F3& v) { return toLocal(makeFrameFromNormal(n), v); }
static inline F3 toWorldFromNormal(const F3& n, const F3& v) { return toWorld(makeFrameFromNormal(n), v); }

// ---------------------------------------------------------------------------
// RNG — include/rt_rng.cuh:8-46
// ---------------------------------------------------------------------------
static inline uint32_t hashUint(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
struct Rng { uint32_t state; };
static inline Rng makeRng(uint32_t seed) { return Rng{seed ? seed : 1u}; }
static inline uint32_t rngNextU32(Rng& r) {
    uint32_t x = r.state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    r.state = x;
    return x;
}
static inline float rngNextFloat01(Rng& r) {
    const uint32_t u = rngNextU32(r) >> 8;
    return (float)u * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// BSDFs — include/rt_bsdf.cuh and per-material headers
// ---------------------------------------------------------------------------
static constexpr float RT_PI = 3.14159265358979323846f;
static constexpr float RT_INV_PI = 1.0f / RT_PI;
static constexpr float AZIMUTH = 1.0f / (2.0f * RT_PI);           // rt_bsdf_microfacet.cuh:11
static constexpr float RT_EPSILON = 1e-4f;
enum { BSDF_Diffuse = 1, BSDF_Dielectric = 2, BSDF_Mirror = 3, BSDF_Microfacet = 4, BSDF_Disney = 5 };
enum { BSDF_ESolidAngle = 1 };

struct BRec {                                                       // rt_bsdf.cuh:18-24
    F3 wi{0, 0, 0};
    F3 wo{0, 0, 0};
    int measure = 0;
    float eta = 0.0f;
    float rndExtra = 0.0f;
};

static inline F3 squareToCosineHemisphere(float u1, float u2) {     // rt_bsdf_diffuse.cuh:11-19
    const float r = std::sqrt(std::fmax(u1, 0.0f));
    const float phi = 2.0f * RT_PI * u2;
    float s, c;
    om_sincos(phi, &s, &c);
    const float x = r * c;
    const float y = r * s;
    const float z = std::sqrt(std::fmax(0.0f, 1.0f - u1));
    return {x, y, z};
}

// ---- diffuse (rt_bsdf_diffuse.cuh:21-62)
static inline F3 evalDiffuse(const OraBsdf& b, const BRec& r) {
    if (r.measure != BSDF_ESolidAngle || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    return mul3(F3{b.p0.x, b.p0.y, b.p0.z}, RT_INV_PI);
}
static inline float pdfDiffuse(const BRec& r) {
    if (r.measure != BSDF_ESolidAngle || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    return RT_INV_PI * r.wo.z;
}
static inline F3 sampleDiffuse(const OraBsdf& b, BRec& r, float u1, float u2) {
    if (r.wi.z <= 0.0f) return {0, 0, 0};
    r.measure = BSDF_ESolidAngle;
    r.wo = squareToCosineHemisphere(u1, u2);
    r.eta = 1.0f;
    return F3{b.p0.x, b.p0.y, b.p0.z};
}

// ---- mirror (rt_bsdf_mirror.cuh:14-33)
static inline F3 sampleMirror(const OraBsdf& b, BRec& r) {
    r.wo = {-r.wi.x, -r.wi.y, r.wi.z};
    r.eta = 1.0f;
    r.measure = BSDF_ESolidAngle;
    return F3{b.p0.x, b.p0.y, b.p0.z};
}

// ---- dielectric (rt_bsdf_dielectric.cuh:13-103)
static inline float fresnelDielectric(float cosThetaI, float etaI, float etaT) {
    if (cosThetaI < 0.0f) {
        const float tmp = etaI; etaI = etaT; etaT = tmp;
        cosThetaI = -cosThetaI;
    }
    const float sinThetaT = (etaI / etaT) * std::sqrt(std::fmax(0.0f, 1.0f - cosThetaI * cosThetaI));
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = std::sqrt(std::fmax(0.0f, 1.0f - sinThetaT * sinThetaT));
    const float Rs = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    const float Rp = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    return (Rs * Rs + Rp * Rp) * 0.5f;
}
static inline F3 sampleDielectric(const OraBsdf& b, BRec& r, float u1) {
    r.measure = BSDF_ESolidAngle;
    F3 result = {1.0f, 1.0f, 1.0f};
    float cosThetaI = r.wi.z;
    const float intIOR = b.p1.x, extIOR = b.p1.y;
    const float fr = fresnelDielectric(cosThetaI, extIOR, intIOR);
    if (u1 < fr) {
        r.wo = {-r.wi.x, -r.wi.y, r.wi.z};
        r.eta = 1.0f;
    } else {
        float indexRatio;
        F3 normal = {0.0f, 0.0f, 1.0f};
        if (cosThetaI > 0) {
            indexRatio = extIOR / intIOR;
        } else {
            indexRatio = intIOR / extIOR;
            normal = {0.0f, 0.0f, -1.0f};
        }
        cosThetaI = std::fabs(cosThetaI);
        const float rootTerm = 1.0f - (indexRatio * indexRatio) * (1.0f - cosThetaI * cosThetaI);
        const float diaScale = indexRatio * cosThetaI - std::sqrt(rootTerm);
        r.wo = {-indexRatio * r.wi.x + diaScale * normal.x,
                -indexRatio * r.wi.y + diaScale * normal.y,
                -indexRatio * r.wi.z + diaScale * normal.z};
        r.eta = 1.0f / indexRatio;
        const float w = indexRatio * indexRatio;
        result = {w, w, w};
    }
    return result;
}

// ---- microfacet (rt_bsdf_microfacet.cuh:23-185)
static inline float tanThetaMf(const F3& w) {
    const float ct = w.z;
    const float sin2 = std::fmax(0.0f, 1.0f - ct * ct);
    const float cos2 = ct * ct;
    if (cos2 <= 0.0f) return 0.0f;
    return std::sqrt(sin2 / cos2);
}
static inline F3 squareToBeckmann(float u1, float u2, float alpha) {
    const float phi = 2.0f * RT_PI * u2;
    const float tan2Theta = -alpha * alpha * om_log(std::fmax(1.0e-6f, 1.0f - u1));
    const float cosTheta = 1.0f / std::sqrt(1.0f + tan2Theta);
    const float sinTheta = std::sqrt(std::fmax(0.0f, 1.0f - cosTheta * cosTheta));
    float s, c;
    om_sincos(phi, &s, &c);
    return {sinTheta * c, sinTheta * s, cosTheta};
}
static inline float squareToBeckmannPdf(const F3& m, float alpha) {
    if (m.z <= 0.0f) return 0.0f;
    const float tanTheta = tanThetaMf(m);
    const float alpha2 = alpha * alpha;
    const float cosTheta = m.z;
    const float cosTheta3 = cosTheta * cosTheta * cosTheta;
    return AZIMUTH * (2.0f * om_exp(-1.0f * tanTheta * tanTheta / alpha2)) / (alpha2 * cosTheta3);
}
static inline float shadowMaskMf(float alpha, const F3& wv, const F3& wh) {
    const float c = dot3(wv, wh) / wv.z;
    if (c <= 0.0f) return 0.0f;
    const float b = 1.0f / (alpha * tanThetaMf(wv));
    if (b < 1.6f) return ((3.535f * b) + 2.181f * b * b) / (1.0f + (2.276f * b) + 2.577f * b * b);
    return 1.0f;
}
static inline F3 evalMicrofacet(const OraBsdf& b, const BRec& r) {
    if (r.measure != BSDF_ESolidAngle || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const float alpha = b.p1.x, intIOR = b.p1.y, extIOR = b.p1.z, ks = b.p1.w;
    const F3 diffuseTerm = {b.p0.x * RT_INV_PI, b.p0.y * RT_INV_PI, b.p0.z * RT_INV_PI};
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float tanValH = tanThetaMf(wh);
    const float cosValH = wh.z;
    const float num = 2.0f * om_exp(-1.0f * tanValH * tanValH / (alpha * alpha));
    const float den = alpha * alpha * cosValH * cosValH * cosValH;
    const float D = AZIMUTH * num / den;
    const float F = fresnelDielectric(dot3(wh, r.wi), extIOR, intIOR);
    const float G = shadowMaskMf(alpha, r.wi, wh) * shadowMaskMf(alpha, r.wo, wh);
    const float cosValI = r.wi.z;
    const float cosValO = r.wo.z;
    const float ecf = 4.0f * cosValI * cosValH * cosValO;
    const float spec = ks * (D * F * G) / ecf;
    return {diffuseTerm.x + spec, diffuseTerm.y + spec, diffuseTerm.z + spec};
}
static inline float pdfMicrofacet(const OraBsdf& b, const BRec& r) {
    if (r.measure != BSDF_ESolidAngle || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const float alpha = b.p1.x, ks = b.p1.w;
    const float cosHemPdf = (r.wo.z <= 0.0f) ? 0.0f : RT_INV_PI * r.wo.z;
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float beckmannPdf = squareToBeckmannPdf(wh, alpha);
    const float jacobian = 1.0f / (4.0f * dot3(wh, r.wo));
    return ks * beckmannPdf * jacobian + (1.0f - ks) * cosHemPdf;
}
static inline F3 sampleMicrofacet(const OraBsdf& b, BRec& r, float u1, float u2) {
    if (r.wi.z <= 0.0f) return {0, 0, 0};
    const float ks = b.p1.w;
    float sampleX = u1;
    int reflectionType;
    if (sampleX < ks) {
        reflectionType = 0;
        sampleX = sampleX / ks;
    } else {
        reflectionType = 1;
        sampleX = (sampleX - ks) / (1.0f - ks);
    }
    r.measure = BSDF_ESolidAngle;
    r.eta = 1.0f;
    if (reflectionType == 0) {
        const F3 n = squareToBeckmann(sampleX, u2, b.p1.x);
        r.wo = sub3(mul3(n, 2.0f * dot3(r.wi, n)), r.wi);
    } else {
        r.wo = squareToCosineHemisphere(sampleX, u2);
    }
    if (r.wo.z <= 0.0f) return {0, 0, 0};
    const float p = pdfMicrofacet(b, r);
    if (p <= 0.0f) return {0, 0, 0};
    return mul3(evalMicrofacet(b, r), r.wo.z / p);
}

// ORA_PLANT (test infrastructure): a nonzero value builds a copy of the oracle
// with one deliberate mistake (1 sheen x2, 2 subsurface clamp dropped, 3 eta^2
// dropped from the rough-glass Jacobian, 4 emitter u0 not rescaled, 5 linear
// spotlight falloff), so the tests can show that their closed-form pins fail
// on it.  The oracle proper is built with ORA_PLANT = 0.
#ifndef ORA_PLANT
#define ORA_PLANT 0
#endif

// ---- Disney (rt_bsdf_disney.cuh:16-892)
// p0.xyz base color, p0.w roughness; p1 metallic, specular, specTrans, specTint;
// p2 sheen, sheenTint, subsurface, anisotropic; p3 clearcoat, clearcoatGloss, eta.
static inline F3 dBase(const OraBsdf& b) { return {b.p0.x, b.p0.y, b.p0.z}; }
static inline F3 sqrt3(const F3& c) {
    return {std::sqrt(std::fmax(0.0f, c.x)), std::sqrt(std::fmax(0.0f, c.y)), std::sqrt(std::fmax(0.0f, c.z))};
}
static inline float diffuseFresnel(float cosTheta, float ninety) {  // :47-49, :55-57
    return 1.0f + ((ninety - 1.0f) * pow5(1.0f - cosTheta));
}
static inline F3 evalDiffuseLobe(const OraBsdf& b, const BRec& r) {  // :60-102
    const float ci = r.wi.z, co = r.wo.z;
    if (ci <= 0.0f || co <= 0.0f) return {0, 0, 0};
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    const float rough = b.p0.w;
    const float dfNinety = 0.5f + (2.0f * rough * hDotWo * hDotWo);
    const float ssNinety = rough * hDotWo * hDotWo;
    const F3 colorOverPi = div3(dBase(b), RT_PI);
    const F3 baseDiffuse = mul3(colorOverPi, diffuseFresnel(ci, dfNinety) * diffuseFresnel(co, dfNinety) * co);
#if ORA_PLANT == 2
    const float ssDenom = ci + co;                      // planted: clamp dropped
#else
    const float ssDenom = std::fmax(1e-2f, ci + co);
#endif
    const F3 subsurface = mul3(colorOverPi,
        1.25f * (diffuseFresnel(ci, ssNinety) * diffuseFresnel(co, ssNinety) * ((1.0f / ssDenom) - 0.5f) + 0.5f) * co);
    const float ss = b.p2.z;
    return add3(mul3(baseDiffuse, 1.0f - ss), mul3(subsurface, ss));
}
static inline float pdfDiffuseLobe(const BRec& r) {                 // :105-121
    if (r.measure != BSDF_ESolidAngle || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
    return RT_INV_PI * r.wo.z;
}
static inline void sampleDiffuseLobe(BRec& r, float u1, float u2) { // :124-143 (value unused)
    if (r.wi.z <= 0) return;
    r.measure = BSDF_ESolidAngle;
    r.wo = squareToCosineHemisphere(u1, u2);
    r.eta = 1.0f;
}
static inline float schlickFresnel(float eta) {
    const float n = eta - 1.0f, d = eta + 1.0f;
    return (n * n) / (d * d);
}
static inline F3 fresnelMetal(const OraBsdf& b, float hDotWo) {     // :154-180
    const F3 base = dBase(b);
    const float lumi = 0.3f * base.x + 0.6f * base.y + 0.1f * base.z;
    F3 tint = {1.0f, 1.0f, 1.0f};
    if (lumi > 0) tint = div3(base, lumi);
    const float specTint = b.p1.w;
    const F3 spec = {(1.0f - specTint) + (specTint * tint.x),
                     (1.0f - specTint) + (specTint * tint.y),
                     (1.0f - specTint) + (specTint * tint.z)};
    const float metallic = b.p1.x;
    const float coef = b.p1.y * schlickFresnel(b.p3.z) * (1.0f - metallic);
    const F3 clr = {(coef * spec.x) + (base.x * metallic),
                    (coef * spec.y) + (base.y * metallic),
                    (coef * spec.z) + (base.z * metallic)};
    const float st = pow5(1.0f - hDotWo);
    return {clr.x + ((1.0f - clr.x) * st), clr.y + ((1.0f - clr.y) * st), clr.z + ((1.0f - clr.z) * st)};
}
static inline float ggxAnisoD(const F3& wh, float ax, float ay) {   // :182-188
    const float q = ((wh.x * wh.x) / (ax * ax)) + ((wh.y * wh.y) / (ay * ay)) + (wh.z * wh.z);
    const float denominator = RT_PI * ax * ay * (q * q);
    return 1.0f / denominator;
}
static inline float smithG1(const F3& w, float ax, float ay) {       // :190-201
    const float axw = w.x * ax;
    const float ayw = w.y * ay;
    const float numerator = std::sqrt(1 + (((axw * axw) + (ayw * ayw)) / std::fmax(1e-8f, (w.z * w.z)))) - 1.0f;
    const float lambda = numerator / 2.0f;
    return 1.0f / (1.0f + lambda);
}
static inline void disneyAlpha(const OraBsdf& b, float& ax, float& ay) {  // :226-231
    const float aspect = std::sqrt(1.0f - (0.9f * b.p2.w));
    const float r2 = b.p0.w * b.p0.w;
    ax = std::fmax(0.0001f, r2 / aspect);
    ay = std::fmax(0.0001f, r2 * aspect);
}
static inline F3 evalMetalLobe(const OraBsdf& b, const BRec& r) {    // :210-241
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    float ax, ay;
    disneyAlpha(b, ax, ay);
    const F3 fm = fresnelMetal(b, hDotWo);
    const float dm = ggxAnisoD(wh, ax, ay);
    const float gm = smithG1(r.wi, ax, ay) * smithG1(r.wo, ax, ay);
    return div3(mul3(fm, dm * gm), (4 * ci));
}
static inline float pdfMetalLobe(const OraBsdf& b, const BRec& r) {  // :244-272
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const F3 wh = normalize3(add3(r.wi, r.wo));
    float ax, ay;
    disneyAlpha(b, ax, ay);
    const float dm = ggxAnisoD(wh, ax, ay);
    const float gm = smithG1(r.wi, ax, ay);
    const float hDotWi = std::fabs(dot3(wh, r.wi));
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    return (dm * gm * hDotWi) / (ci * 4.0f * hDotWo);
}
// Heitz visible-normal sampling, shared by the metal and glass lobes (:290-318, :597-626).
static inline F3 sampleVisibleNormal(const F3& wUpper, float ax, float ay, float u1, float u2) {
    const F3 vh = normalize3(F3{ax * wUpper.x, ay * wUpper.y, wUpper.z});
    const float pl = (vh.x * vh.x) + (vh.y * vh.y);
    F3 orthOne = {1.0f, 0.0f, 0.0f};
    if (pl > 0.0f) orthOne = div3(F3{-1.0f * vh.y, vh.x, 0.0f}, std::sqrt(pl));
    const F3 orthTwo = cross3(vh, orthOne);
    const float radius = std::sqrt(u1);
    const float phi = 2.0f * RT_PI * u2;
    float s, c;
    om_sincos(phi, &s, &c);
    const float t1 = radius * c;
    float t2 = radius * s;
    const float skew = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - skew) * std::sqrt(1.0f - (t1 * t1)) + (skew * t2);
    const F3 ph = add3(add3(mul3(orthOne, t1), mul3(orthTwo, t2)),
                       mul3(vh, std::sqrt(std::fmax(0.0f, 1 - (t1 * t1) - (t2 * t2)))));
    return normalize3(F3{ax * ph.x, ay * ph.y, std::fmax(0.0f, ph.z)});
}
static inline void sampleMetalLobe(const OraBsdf& b, BRec& r, float u1, float u2) {  // :275-335
    if (r.wi.z <= 0.0f) return;
    r.measure = BSDF_ESolidAngle;
    r.eta = 1.0f;
    float ax, ay;
    disneyAlpha(b, ax, ay);
    const F3 randH = sampleVisibleNormal(r.wi, ax, ay, u1, u2);
    r.wo = sub3(mul3(randH, 2.0f * dot3(r.wi, randH)), r.wi);
}
static inline float fresnelClearcoat(float hDotWo) {                // :339-349
    const float rr = 0.04f;
    return rr + ((1.0f - rr) * pow5(1.0f - hDotWo));
}
static inline float clearcoatIr2(const OraBsdf& b) {
    const float g = b.p3.y;
    const float ir = ((1.0f - g) * 0.1f) + (g * 0.001f);
    return ir * ir;
}
static inline float distributionClearcoat(const OraBsdf& b, const F3& wh) {  // :351-362
    const float irSquared = clearcoatIr2(b);
    const float numerator = irSquared - 1.0f;
    const float denominator = RT_PI * om_log(irSquared) * (1.0f + ((irSquared - 1.0f) * (wh.z * wh.z)));
    return numerator / denominator;
}
static inline F3 evalClearcoatLobe(const OraBsdf& b, const BRec& r) {  // :369-394
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    const float fc = fresnelClearcoat(hDotWo);
    const float dc = distributionClearcoat(b, wh);
    const float gc = smithG1(r.wi, 0.25f, 0.25f) * smithG1(r.wo, 0.25f, 0.25f);
    const float denominator = 4.0f * ci;
    const float cc = (fc * dc * gc) / denominator;
    return {cc, cc, cc};
}
static inline float pdfClearcoatLobe(const OraBsdf& b, const BRec& r) {  // :397-415
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float dc = distributionClearcoat(b, wh);
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    return (dc * wh.z) / (4.0f * hDotWo);
}
static inline void sampleClearcoatLobe(const OraBsdf& b, BRec& r, float u1, float u2) {  // :418-451
    if (r.wi.z <= 0.0f) return;
    r.measure = BSDF_ESolidAngle;
    const float irSquared = clearcoatIr2(b);
    const float azimuth = 2.0f * RT_PI * u1;
    const float cosElevationSquared = (1.0f - om_pow(irSquared, 1.0f - u2)) / (1.0f - irSquared);
    const float sinElevation = std::sqrt(1.0f - cosElevationSquared);
    const float hz = std::sqrt(cosElevationSquared);
    float s, c;
    om_sincos(azimuth, &s, &c);
    const float hy = sinElevation * s;
    const float hx = sinElevation * c;
    const F3 wh = normalize3(F3{hx, hy, hz});
    r.wo = sub3(mul3(wh, 2.0f * dot3(r.wi, wh)), r.wi);
    r.eta = 1.0f;
}
static inline float fresnelGlass(float hDotWi, float hDotWo, float etahDotWi, float etahDotWo) {  // :455-472
    hDotWi = std::fabs(hDotWi);
    hDotWo = std::fabs(hDotWo);
    etahDotWi = std::fabs(etahDotWi);
    etahDotWo = std::fabs(etahDotWo);
    const float rs = (hDotWi - etahDotWo) / (hDotWi + etahDotWo);
    const float rp = (etahDotWi - hDotWo) / (etahDotWi + hDotWo);
    return 0.5f * ((rs * rs) + (rp * rp));
}
static inline F3 glassHalfVector(const BRec& r, float eta, float inNout) {
    F3 wh;
    if (inNout > 0.0f) wh = normalize3(add3(r.wi, r.wo));
    else wh = mul3(normalize3(add3(r.wi, mul3(r.wo, eta))), -1.0f);
    if (wh.z < 0.0f) wh = mul3(wh, -1.0f);
    return wh;
}
static inline F3 evalGlassLobe(const OraBsdf& b, const BRec& r) {   // :475-533
    const float ci = r.wi.z, co = r.wo.z;
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const float inNout = ci * co;
    const F3 wh = glassHalfVector(r, eta, inNout);
    const float hDotWi = dot3(wh, r.wi);
    const float hDotWo = dot3(wh, r.wo);
    const float etahDotWi = eta * hDotWi;
    const float etahDotWo = eta * hDotWo;
    const float fg = fresnelGlass(hDotWi, hDotWo, etahDotWi, etahDotWo);
    float ax, ay;
    disneyAlpha(b, ax, ay);
    const float dg = ggxAnisoD(wh, ax, ay);
    const float gg = smithG1(r.wi, ax, ay) * smithG1(r.wo, ax, ay);
    if (inNout > 0.0f) {
        const F3 numerator = mul3(dBase(b), fg * dg * gg);
        const float denominator = 4.0f * std::fabs(ci);
        return div3(numerator, denominator);
    }
    const F3 numerator = mul3(sqrt3(dBase(b)), (1.0f - fg) * dg * gg * std::fabs(hDotWi * hDotWo));
    const float q = hDotWi + etahDotWo;
    const float denominator = std::fabs(ci) * std::fmax(q * q, 1e-4f);
    return div3(numerator, denominator);
}
static inline float pdfGlassLobe(const OraBsdf& b, const BRec& r) {  // :536-583
    const float ci = r.wi.z, co = r.wo.z;
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const float inNout = ci * co;
    const F3 wh = glassHalfVector(r, eta, inNout);
    const float hDotWi = dot3(wh, r.wi);
    const float hDotWo = dot3(wh, r.wo);
    const float fg = fresnelGlass(hDotWi, hDotWo, eta * hDotWi, eta * hDotWo);
    float ax, ay;
    disneyAlpha(b, ax, ay);
    const float dg = ggxAnisoD(wh, ax, ay);
    const float gg = smithG1(r.wi, ax, ay);
    const float commonPDF = (dg * gg * std::fabs(hDotWi)) / std::fabs(ci);
    if (inNout > 0.0f) return (commonPDF * fg) / (4.0f * std::fabs(hDotWo));
    const float q = hDotWi + (eta * hDotWo);
#if ORA_PLANT == 3
    const float jacobian = std::fabs(hDotWo) / std::fmax(q * q, 1e-4f);   // planted: eta^2 dropped
#else
    const float jacobian = (eta * eta * std::fabs(hDotWo)) / std::fmax(q * q, 1e-4f);
#endif
    return (1.0f - fg) * commonPDF * jacobian;
}
static inline void sampleGlassLobe(const OraBsdf& b, BRec& r, float u1, float u2) {  // :586-674
    const float ci = r.wi.z;
    r.measure = BSDF_ESolidAngle;
    float ax, ay;
    disneyAlpha(b, ax, ay);
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const F3 wUpper = (ci < 0.0f) ? mul3(r.wi, -1.0f) : r.wi;
    F3 randH = sampleVisibleNormal(wUpper, ax, ay, u1, u2);
    if (ci < 0.0f) randH = mul3(randH, -1.0f);
    const float hDotWi = dot3(randH, r.wi);
    const float sinSqThetaT = (1.0f - hDotWi * hDotWi) / (eta * eta);
    const float hDotWt = std::sqrt(std::fmax(0.0f, 1.0f - sinSqThetaT));
    const float fg = fresnelGlass(hDotWi, hDotWt, eta * hDotWi, eta * hDotWt);
    if (r.rndExtra < fg || sinSqThetaT >= 1.0f) {
        r.wo = sub3(mul3(randH, 2.0f * dot3(r.wi, randH)), r.wi);
        r.eta = 1.0f;
    } else {
        r.wo = sub3(mul3(randH, (hDotWi / eta) - hDotWt), div3(r.wi, eta));
        r.wo = normalize3(r.wo);
        r.eta = eta;
    }
}
static inline F3 evalSheenLobe(const OraBsdf& b, const BRec& r) {   // :679-711
    const float ci = r.wi.z, co = r.wo.z;
    if (ci <= 0.0f || co <= 0.0f) return {0, 0, 0};
    const F3 wh = normalize3(add3(r.wi, r.wo));
    const float hDotWo = std::fabs(dot3(wh, r.wo));
    const F3 base = dBase(b);
    const float lumi = 0.3f * base.x + 0.6f * base.y + 0.1f * base.z;
    F3 tint = {1.0f, 1.0f, 1.0f};
    if (lumi > 0) tint = div3(base, lumi);
    const float sheenTint = b.p2.y;
    const F3 sheen = {(1.0f - sheenTint) + (sheenTint * tint.x),
                      (1.0f - sheenTint) + (sheenTint * tint.y),
                      (1.0f - sheenTint) + (sheenTint * tint.z)};
#if ORA_PLANT == 1
    return mul3(sheen, 2.0f * pow5(1.0f - hDotWo) * co);   // planted: sheen x2
#else
    return mul3(sheen, pow5(1.0f - hDotWo) * co);
#endif
}
static inline F3 evalDisney(const OraBsdf& b, const BRec& r) {      // :758-787
    const float specT = b.p1.z;
    if (r.wi.z <= 0.0f) {
        if (specT <= 0.0f) return {0, 0, 0};
        return evalGlassLobe(b, r);
    }
    const float minMetal = (1.0f - b.p1.x);
    const float minSpecTransmission = (1.0f - specT);
    const F3 fd = mul3(evalDiffuseLobe(b, r), minMetal * minSpecTransmission);
    const F3 fs = mul3(evalSheenLobe(b, r), minMetal * b.p2.x);
    const F3 fm = mul3(evalMetalLobe(b, r), (1.0f - (specT * minMetal)));
    const F3 fc = mul3(evalClearcoatLobe(b, r), 0.25f * b.p3.x);
    const F3 fg = mul3(evalGlassLobe(b, r), minMetal * specT);
    return add3(add3(add3(add3(fd, fm), fg), fc), fs);
}
static inline float pdfDisney(const OraBsdf& b, const BRec& r) {    // :790-820
    const float specT = b.p1.z;
    const float minMetal = (1.0f - b.p1.x);
    const float minSpecTransmission = (1.0f - specT);
    const float diffuseWeight = minMetal * minSpecTransmission;
    const float metallicWeight = 1.0f - (specT * minMetal);
    const float glassWeight = minMetal * specT;
    const float clearcoatWeight = 0.25f * b.p3.x;
    const float weightSum = diffuseWeight + glassWeight + metallicWeight + clearcoatWeight;
    if (r.wi.z <= 0) {
        if (specT <= 0.0f) return 0.0f;
        return pdfGlassLobe(b, r);
    }
    const float probDiffuse = pdfDiffuseLobe(r);
    const float probMetal = pdfMetalLobe(b, r);
    const float probGlass = pdfGlassLobe(b, r);
    const float probClearcoat = pdfClearcoatLobe(b, r);
    return (diffuseWeight * probDiffuse + glassWeight * probGlass + metallicWeight * probMetal +
            clearcoatWeight * probClearcoat) / weightSum;
}
static inline F3 sampleDisney(const OraBsdf& b, BRec& r, float u1, float u2, float* outPdf) {  // :823-892
    const float specT = b.p1.z;
    if (r.wi.z <= 0.0f) {
        if (specT <= 0.0f) {
            if (outPdf) *outPdf = 0.0f;
            return {0, 0, 0};
        }
        sampleGlassLobe(b, r, u1, u2);
    } else {
        const float minMetal = (1.0f - b.p1.x);
        const float diffuseWeight = minMetal * (1.0f - specT);
        const float specularWeight = minMetal * specT;
        const float metallicWeight = 1.0f - (specT * minMetal);
        const float clearcoatWeight = 0.25f * b.p3.x;
        const float weightSum = diffuseWeight + specularWeight + metallicWeight + clearcoatWeight;
        const float rndChoice = r.rndExtra;
        const float diffuseFrac = diffuseWeight / weightSum;
        const float specularFrac = specularWeight / weightSum;
        const float metallicFrac = metallicWeight / weightSum;
        if (rndChoice < diffuseFrac) {
            sampleDiffuseLobe(r, u1, u2);
        } else if (rndChoice < diffuseFrac + specularFrac) {
            r.rndExtra = (specularFrac > 0.0f) ? (rndChoice - diffuseFrac) / specularFrac : 0.0f;
            sampleGlassLobe(b, r, u1, u2);
        } else if (rndChoice < diffuseFrac + specularFrac + metallicFrac) {
            sampleMetalLobe(b, r, u1, u2);
        } else {
            sampleClearcoatLobe(b, r, u1, u2);
        }
    }
    const float disneyPDF = pdfDisney(b, r);
    if (outPdf) *outPdf = disneyPDF;
    if (disneyPDF < 1e-7f) return {0, 0, 0};
    return div3(evalDisney(b, r), disneyPDF);
}

// ---- dispatch (rt_bsdf.cuh:42-84)
static inline F3 bsdfEval(const OraBsdf& b, const BRec& r) {
    switch (b.type) {
        case BSDF_Diffuse: return evalDiffuse(b, r);
        case BSDF_Microfacet: return evalMicrofacet(b, r);
        case BSDF_Disney: return evalDisney(b, r);
        default: return {0, 0, 0};    // dielectric, mirror: delta
    }
}
static inline float bsdfPdf(const OraBsdf& b, const BRec& r) {
    switch (b.type) {
        case BSDF_Diffuse: return pdfDiffuse(r);
        case BSDF_Microfacet: return pdfMicrofacet(b, r);
        case BSDF_Disney: return pdfDisney(b, r);
        default: return 0.0f;
    }
}
static inline F3 bsdfSample(const OraBsdf& b, BRec& r, float u1, float u2) {
    switch (b.type) {
        case BSDF_Diffuse: return sampleDiffuse(b, r, u1, u2);
        case BSDF_Dielectric: return sampleDielectric(b, r, u1);
        case BSDF_Mirror: return sampleMirror(b, r);
        case BSDF_Microfacet: return sampleMicrofacet(b, r, u1, u2);
        case BSDF_Disney: return sampleDisney(b, r, u1, u2, nullptr);
        default: return {0, 0, 0};
    }
}
static inline bool bsdfIsDiffuse(const OraBsdf& b) {
    return b.type == BSDF_Diffuse || b.type == BSDF_Microfacet || b.type == BSDF_Disney;
}

// ---------------------------------------------------------------------------
// Scene, BVH (oracle's own binned-SAH BVH2) and watertight intersection
// ---------------------------------------------------------------------------
struct Node { float lo[3], hi[3]; int left, right, start, count; };

struct Scene {
    OraSceneDesc d;
    std::vector<Node> nodes;
    std::vector<int> triOrder;
    std::vector<int> emitterOfTri;      // first emitter listing each triangle, or -1
    bool hasEnv = false;
    float envRot = 0.0f;
};

// Binned surface-area-heuristic build (16 bins per axis on the centroid
// bounds; cost 1 per node visit plus 1 per triangle test, weighted by child
// area): a tree of the quality a production builder gives, so the oracle's
// traversal is a fair CPU baseline.  Any tree gives the same hits (the
// triangle test clamps to the triangle's own box), so the shape only affects
// speed.  Centroids that coincide fall back to an object-median split.
constexpr int kSahBins = 16;
constexpr int kMaxLeaf = 8;

static float boxArea(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

static void makeLeaf(Node& nd, int start, int count) {
    nd.left = nd.right = -1;
    nd.start = start;
    nd.count = count;
}

static void buildNode(Scene& s, const std::vector<float>& cent, const std::vector<float>& tbox, int nodeIdx,
                      int start, int count, int depth) {
    {
        Node& nd = s.nodes[nodeIdx];
        for (int k = 0; k < 3; ++k) { nd.lo[k] = INFINITY; nd.hi[k] = -INFINITY; }
        for (int i = start; i < start + count; ++i) {
            const float* b = &tbox[6 * (size_t)s.triOrder[i]];
            for (int k = 0; k < 3; ++k) { nd.lo[k] = std::min(nd.lo[k], b[k]); nd.hi[k] = std::max(nd.hi[k], b[3 + k]); }
        }
        if (count <= 2 || depth > 60) {
            makeLeaf(nd, start, count);
            return;
        }
    }
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = start; i < start + count; ++i)
        for (int k = 0; k < 3; ++k) {
            clo[k] = std::min(clo[k], cent[3 * (size_t)s.triOrder[i] + k]);
            chi[k] = std::max(chi[k], cent[3 * (size_t)s.triOrder[i] + k]);
        }
    const float parentArea = boxArea(s.nodes[nodeIdx].lo, s.nodes[nodeIdx].hi);
    float bestCost = INFINITY;
    int bestAxis = -1, bestSplit = 0;
    for (int k = 0; k < 3; ++k) {
        const float ext = chi[k] - clo[k];
        if (!(ext > 0.0f)) continue;
        const float scale = (float)kSahBins / ext;
        int cnt[kSahBins] = {};
        float blo[kSahBins][3], bhi[kSahBins][3];
        for (int b = 0; b < kSahBins; ++b)
            for (int j = 0; j < 3; ++j) { blo[b][j] = INFINITY; bhi[b][j] = -INFINITY; }
        for (int i = start; i < start + count; ++i) {
            const int id = s.triOrder[i];
            const int b = std::min(kSahBins - 1, (int)((cent[3 * (size_t)id + k] - clo[k]) * scale));
            ++cnt[b];
            const float* tb = &tbox[6 * (size_t)id];
            for (int j = 0; j < 3; ++j) { blo[b][j] = std::min(blo[b][j], tb[j]); bhi[b][j] = std::max(bhi[b][j], tb[3 + j]); }
        }
        float rArea[kSahBins];
        int rCnt[kSahBins];
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int n = 0;
        for (int b = kSahBins - 1; b > 0; --b) {
            n += cnt[b];
            for (int j = 0; j < 3; ++j) { lo[j] = std::min(lo[j], blo[b][j]); hi[j] = std::max(hi[j], bhi[b][j]); }
            rCnt[b] = n;
            rArea[b] = n ? boxArea(lo, hi) : 0.0f;
        }
        for (int j = 0; j < 3; ++j) { lo[j] = INFINITY; hi[j] = -INFINITY; }
        n = 0;
        for (int b = 0; b < kSahBins - 1; ++b) {      // split between bins b and b + 1
            n += cnt[b];
            for (int j = 0; j < 3; ++j) { lo[j] = std::min(lo[j], blo[b][j]); hi[j] = std::max(hi[j], bhi[b][j]); }
            if (n == 0 || rCnt[b + 1] == 0) continue;
            const float cost = 1.0f + (boxArea(lo, hi) * (float)n + rArea[b + 1] * (float)rCnt[b + 1]) / parentArea;
            if (cost < bestCost) { bestCost = cost; bestAxis = k; bestSplit = b + 1; }
        }
    }
    int mid;
    if (bestAxis >= 0) {
        if (count <= kMaxLeaf && (float)count <= bestCost) {
            makeLeaf(s.nodes[nodeIdx], start, count);
            return;
        }
        const float scale = (float)kSahBins / (chi[bestAxis] - clo[bestAxis]);
        const float c0 = clo[bestAxis];
        mid = (int)(std::stable_partition(s.triOrder.begin() + start, s.triOrder.begin() + start + count,
                                          [&](int id) {
                                              return std::min(kSahBins - 1,
                                                              (int)((cent[3 * (size_t)id + bestAxis] - c0) * scale)) <
                                                     bestSplit;
                                          }) -
                    s.triOrder.begin());
    } else {
        // coincident centroids: object median on index order
        if (count <= kMaxLeaf) {
            makeLeaf(s.nodes[nodeIdx], start, count);
            return;
        }
        mid = start + count / 2;
    }
    const int l = (int)s.nodes.size();
    s.nodes.push_back(Node{});
    s.nodes.push_back(Node{});
    s.nodes[nodeIdx].left = l; s.nodes[nodeIdx].right = l + 1; s.nodes[nodeIdx].start = 0; s.nodes[nodeIdx].count = 0;
    buildNode(s, cent, tbox, l, start, mid - start, depth + 1);
    buildNode(s, cent, tbox, l + 1, mid, start + count - mid, depth + 1);
}

static void buildBvh(Scene& s) {
    const int T = s.d.triangleCount;
    s.triOrder.resize(T);
    for (int i = 0; i < T; ++i) s.triOrder[i] = i;
    std::vector<float> cent(3 * (size_t)T), tbox(6 * (size_t)T);
    for (int i = 0; i < T; ++i) {
        const OraTriangle& t = s.d.triangles[i];
        cent[3 * i + 0] = (t.v0.x + t.v1.x + t.v2.x) / 3.0f;
        cent[3 * i + 1] = (t.v0.y + t.v1.y + t.v2.y) / 3.0f;
        cent[3 * i + 2] = (t.v0.z + t.v1.z + t.v2.z) / 3.0f;
        const F3 v[3] = {t.v0, t.v1, t.v2};
        float* b = &tbox[6 * (size_t)i];
        b[0] = b[1] = b[2] = INFINITY;
        b[3] = b[4] = b[5] = -INFINITY;
        for (int j = 0; j < 3; ++j) {
            const float c[3] = {v[j].x, v[j].y, v[j].z};
            for (int k = 0; k < 3; ++k) { b[k] = std::min(b[k], c[k]); b[3 + k] = std::max(b[3 + k], c[k]); }
        }
    }
    s.nodes.clear();
    s.nodes.reserve(2 * (size_t)T + 1);
    s.nodes.push_back(Node{});
    if (T > 0) buildNode(s, cent, tbox, 0, 0, T, 0);
    else { s.nodes[0].left = s.nodes[0].right = -1; s.nodes[0].start = 0; s.nodes[0].count = 0;
           for (int k = 0; k < 3; ++k) { s.nodes[0].lo[k] = INFINITY; s.nodes[0].hi[k] = -INFINITY; } }
}

// Per-ray constants of the watertight test (Woop et al. 2013, section 3).
struct WtRay {
    F3 o, d;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float inv[3];
    bool valid;
};
static inline WtRay makeWtRay(const F3& o, const F3& d) {
    WtRay r;
    r.o = o; r.d = d;
    const float ad[3] = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    r.valid = (ad[0] > 0.0f || ad[1] > 0.0f || ad[2] > 0.0f);
    int kz = 0;
    if (ad[1] > ad[kz]) kz = 1;
    if (ad[2] > ad[kz]) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    const float dv[3] = {d.x, d.y, d.z};
    if (dv[kz] < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = dv[kx] / dv[kz];
    r.Sy = dv[ky] / dv[kz];
    r.Sz = 1.0f / dv[kz];
    for (int k = 0; k < 3; ++k) r.inv[k] = 1.0f / dv[k];
    return r;
}
static inline float comp(const F3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Returns true and (t, u, v) if the ray hits the triangle with t in [tmin, tmax];
// u and v are the barycentric weights of v1 and v2.
static inline bool intersectTri(const WtRay& r, const OraTriangle& tri, float tmin, float tmax,
                                float& tOut, float& uOut, float& vOut) {
    const F3 A = sub3(tri.v0, r.o);
    const F3 B = sub3(tri.v1, r.o);
    const F3 C = sub3(tri.v2, r.o);
    const float Az_ = comp(A, r.kz), Bz_ = comp(B, r.kz), Cz_ = comp(C, r.kz);
    const float Ax = comp(A, r.kx) - r.Sx * Az_;
    const float Ay = comp(A, r.ky) - r.Sy * Az_;
    const float Bx = comp(B, r.kx) - r.Sx * Bz_;
    const float By = comp(B, r.ky) - r.Sy * Bz_;
    const float Cx = comp(C, r.kx) - r.Sx * Cz_;
    const float Cy = comp(C, r.ky) - r.Sy * Cz_;
    const float U = Cx * By - Cy * Bx;
    const float V = Ax * Cy - Ay * Cx;
    const float W = Bx * Ay - By * Ax;
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Az_;
    const float Bz = r.Sz * Bz_;
    const float Cz = r.Sz * Cz_;
    const float T = U * Az + V * Bz + W * Cz;
    // T/det is badly conditioned for a ray nearly parallel to a large
    // triangle: the distance is clamped to the line interval of the
    // triangle's own box, computed as hitBox computes a node's (an empty
    // interval is a miss).  Rounding is monotone, so every node box holding
    // the triangle admits the clamped distance and the BVH finds exactly what
    // the loop over all triangles finds.
    float lo = -INFINITY, hi = INFINITY;
    const F3 P[3] = {tri.v0, tri.v1, tri.v2};
    for (int k = 0; k < 3; ++k) {
        const float o = comp(r.o, k);
        const float a = (std::fmin(std::fmin(comp(P[0], k), comp(P[1], k)), comp(P[2], k)) - o) * r.inv[k];
        const float b = (std::fmax(std::fmax(comp(P[0], k), comp(P[1], k)), comp(P[2], k)) - o) * r.inv[k];
        lo = std::fmax(lo, std::fmin(a, b));
        hi = std::fmin(hi, std::fmax(a, b));
    }
    if (!(lo <= hi)) return false;
    const float t = std::fmin(std::fmax(T / det, lo), hi);
    if (!(t >= tmin && t <= tmax)) return false;
    tOut = t;
    uOut = V / det;
    vOut = W / det;
    return true;
}

// Slab test of a node box against [tmin, tmax].  No widening is needed: a
// triangle's hit distance is clamped to its own box interval (intersectTri),
// which any enclosing node's computed interval contains, so the answer is the
// minimum over all triangles whatever the tree (equal to the brute-force loop
// bit for bit).
static inline bool hitBox(const WtRay& r, const Node& n, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    for (int k = 0; k < 3; ++k) {
        const float a = (n.lo[k] - o[k]) * r.inv[k];
        const float b = (n.hi[k] - o[k]) * r.inv[k];
        const float lo = std::fmin(a, b), hi = std::fmax(a, b);
        t0 = std::fmax(t0, lo);
        t1 = std::fmin(t1, hi);
    }
    return t0 <= t1;
}

struct Hit { int tri = -1; float t = 0, u = 0, v = 0; };

// Slab test of a node's box; the entry distance in t0.
static inline bool boxEntry(const WtRay& r, const Node& n, float tmin, float tmax, float& t0) {
    float t1 = tmax;
    t0 = tmin;
    const float o[3] = {r.o.x, r.o.y, r.o.z};
    for (int k = 0; k < 3; ++k) {
        const float a = (n.lo[k] - o[k]) * r.inv[k];
        const float b = (n.hi[k] - o[k]) * r.inv[k];
        const float lo = std::fmin(a, b), hi = std::fmax(a, b);
        t0 = std::fmax(t0, lo);
        t1 = std::fmin(t1, hi);
    }
    return t0 <= t1;
}

// Children are visited nearer entry first; a postponed child is skipped when a
// hit found meanwhile lies before its entry, which is hitBox against the
// current best distance (its slab interval is already known to be non-empty).
static Hit closestHit(const Scene& s, const F3& o, const F3& d, float tmin, float tmax) {
    Hit best;
    const WtRay r = makeWtRay(o, d);
    if (!r.valid || s.d.triangleCount == 0) return best;
    float bestT = tmax;
    int stack[128];
    float entry[128];
    int sp = 0;
    float e0;
    if (!boxEntry(r, s.nodes[0], tmin, bestT, e0)) return best;
    stack[sp] = 0;
    entry[sp++] = e0;
    while (sp) {
        --sp;
        if (entry[sp] > bestT) continue;
        const Node& n = s.nodes[stack[sp]];
        if (n.left < 0) {
            for (int i = n.start; i < n.start + n.count; ++i) {
                const int id = s.triOrder[i];
                float t, u, v;
                if (intersectTri(r, s.d.triangles[id], tmin, bestT, t, u, v)) {
                    if (t < bestT || best.tri < 0 || (t == bestT && id < best.tri)) {
                        bestT = t; best.tri = id; best.t = t; best.u = u; best.v = v;
                    }
                }
            }
            continue;
        }
        float el, er;
        const bool hl = boxEntry(r, s.nodes[n.left], tmin, bestT, el);
        const bool hr = boxEntry(r, s.nodes[n.right], tmin, bestT, er);
        if (hl && hr) {
            const bool leftFirst = el <= er;
            stack[sp] = leftFirst ? n.right : n.left;
            entry[sp++] = leftFirst ? er : el;
            stack[sp] = leftFirst ? n.left : n.right;
            entry[sp++] = leftFirst ? el : er;
        } else if (hl) {
            stack[sp] = n.left;
            entry[sp++] = el;
        } else if (hr) {
            stack[sp] = n.right;
            entry[sp++] = er;
        }
    }
    return best;
}

static bool occluded(const Scene& s, const F3& o, const F3& d, float tmax) {
    const WtRay r = makeWtRay(o, d);
    if (!r.valid || s.d.triangleCount == 0) return false;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& n = s.nodes[stack[--sp]];
        if (!hitBox(r, n, RT_EPSILON, tmax)) continue;
        if (n.left < 0) {
            for (int i = n.start; i < n.start + n.count; ++i) {
                float t, u, v;
                if (intersectTri(r, s.d.triangles[s.triOrder[i]], RT_EPSILON, tmax, t, u, v)) return true;
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return false;
}

static Hit bruteClosest(const Scene& s, const F3& o, const F3& d, float tmin, float tmax) {
    Hit best;
    const WtRay r = makeWtRay(o, d);
    if (!r.valid) return best;
    float bestT = tmax;
    for (int id = 0; id < s.d.triangleCount; ++id) {
        float t, u, v;
        if (intersectTri(r, s.d.triangles[id], tmin, bestT, t, u, v)) {
            if (t < bestT || best.tri < 0) { bestT = t; best.tri = id; best.t = t; best.u = u; best.v = v; }
        }
    }
    return best;
}

// ---------------------------------------------------------------------------
// Env map — include/rt_env_map.cuh (exact bilinear filtering, see header)
// ---------------------------------------------------------------------------
static inline F3 envTexel(const Scene& s, int ix, int iy) {
    const float* p = s.d.envRGBA + 4 * ((size_t)iy * (size_t)s.d.envWidth + (size_t)ix);
    return {p[0], p[1], p[2]};
}
static inline F3 envBilinear(const Scene& s, float u, float v) {
    const int W = s.d.envWidth, H = s.d.envHeight;
    float x = u * (float)W - 0.5f;
    float y = v * (float)H - 0.5f;
    if (!(x == x)) x = 0.0f;
    if (!(y == y)) y = 0.0f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    int ix = (int)fx, iy = (int)fy;
    int ix0 = ix % W; if (ix0 < 0) ix0 += W;
    int ix1 = ix0 + 1; if (ix1 == W) ix1 = 0;
    const int iy0 = std::min(std::max(iy, 0), H - 1);
    const int iy1 = std::min(std::max(iy + 1, 0), H - 1);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    const F3 t00 = envTexel(s, ix0, iy0), t10 = envTexel(s, ix1, iy0), t01 = envTexel(s, ix0, iy1), t11 = envTexel(s, ix1, iy1);
    return {w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
            w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
            w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z};
}
static inline F2 dirToEquirect(const F3& dir, float rot) {         // :36-52
    float phi = om_atan2(dir.z, dir.x) + rot;
    const float theta = om_acos(std::fmin(std::fmax(dir.y, -1.0f), 1.0f));
    phi = std::fmod(phi + F_PI, 2.0f * F_PI);
    if (phi < 0.0f) phi += 2.0f * F_PI;
    phi -= F_PI;
    return {(phi + F_PI) / (2.0f * F_PI), theta / F_PI};
}
static inline F3 equirectToDir(float u, float v, float rot) {      // :56-67
    const float phi = u * 2.0f * F_PI - F_PI - rot;
    const float theta = v * F_PI;
    float st, ct, sp, cp;
    om_sincos(theta, &st, &ct);
    om_sincos(phi, &sp, &cp);
    return normalize3({st * cp, ct, st * sp});
}
static inline F3 envEval(const Scene& s, const F3& dir) {          // :72-77
    const F2 uv = dirToEquirect(dir, s.envRot);
    const F3 t = envBilinear(s, uv.x, uv.y);
    const float k = s.d.envIntensity;
    return {t.x * k, t.y * k, t.z * k};
}
static inline int cdfSearch(const float* cdf, int count, float x) { // :82-91
    int lo = 0, hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}
static inline float envPdf(const Scene& s, const F3& dir) {        // :97-132
    if (s.d.envTotalPower <= 0.0f) return 0.0f;
    const int W = s.d.envWidth, H = s.d.envHeight;
    const F2 uv = dirToEquirect(dir, s.envRot);
    const float fx = uv.x * (float)W;
    const float fy = uv.y * (float)H;
    // (int) of the clamped value: identical to the reference's clamp of (int)fx
    // for every finite fx, and defined for NaN / huge values.
    const int ix = std::min(std::max((int)std::fmin(std::fmax(fx, 0.0f), (float)W), 0), W - 1);
    const int iy = std::min(std::max((int)std::fmin(std::fmax(fy, 0.0f), (float)H), 0), H - 1);
    const float* marg = s.d.envMarginalCdf;
    const float marginalSum = marg[H];
    const float rowWeight = marg[iy + 1] - marg[iy];
    if (marginalSum <= 0.0f || rowWeight <= 0.0f) return 0.0f;
    const float pdfRow = rowWeight / marginalSum;
    const float* rowCdf = s.d.envConditionalCdf + (size_t)iy * (size_t)(W + 1);
    const float rowSum = rowCdf[W];
    const float colWeight = rowCdf[ix + 1] - rowCdf[ix];
    if (rowSum <= 0.0f || colWeight <= 0.0f) return 0.0f;
    const float pdfCol = colWeight / rowSum;
    const float theta = uv.y * F_PI;
    const float sinTheta = std::fmax(om_sin(theta), 1e-6f);
    return (pdfRow * pdfCol * (float)W * (float)H) / (2.0f * F_PI * F_PI * sinTheta);
}
struct EnvSample { F3 direction; float pdf; F3 radiance; };
static inline EnvSample envSample(const Scene& s, float u0, float u1) {  // :144-199
    EnvSample r;
    r.direction = {0.0f, 1.0f, 0.0f};
    r.pdf = 0.0f;
    r.radiance = {0.0f, 0.0f, 0.0f};
    if (s.d.envTotalPower <= 0.0f) return r;
    const int W = s.d.envWidth, H = s.d.envHeight;
    const float* marg = s.d.envMarginalCdf;
    const float marginalSum = marg[H];
    if (marginalSum <= 0.0f) return r;
    const float xiRow = u0 * marginalSum;
    const int row = cdfSearch(marg, H, xiRow);
    const float c0r = marg[row], c1r = marg[row + 1];
    const float dRow = c1r - c0r;
    const float fracRow = (dRow > 0.0f) ? (xiRow - c0r) / dRow : 0.0f;
    const float* rowCdf = s.d.envConditionalCdf + (size_t)row * (size_t)(W + 1);
    const float rowSum = rowCdf[W];
    if (rowSum <= 0.0f) return r;
    const float xiCol = u1 * rowSum;
    const int col = cdfSearch(rowCdf, W, xiCol);
    const float c0c = rowCdf[col], c1c = rowCdf[col + 1];
    const float dCol = c1c - c0c;
    const float fracCol = (dCol > 0.0f) ? (xiCol - c0c) / dCol : 0.0f;
    const float u = ((float)col + fracCol) / (float)W;
    const float v = ((float)row + fracRow) / (float)H;
    r.direction = equirectToDir(u, v, s.envRot);
    const F3 t = envBilinear(s, u, v);
    const float k = s.d.envIntensity;
    r.radiance = {t.x * k, t.y * k, t.z * k};
    const float pdfRow = dRow / marginalSum;
    const float pdfCol = dCol / rowSum;
    const float theta = v * F_PI;
    const float sinTheta = std::fmax(om_sin(theta), 1e-6f);
    r.pdf = (pdfRow * pdfCol * (float)W * (float)H) / (2.0f * F_PI * F_PI * sinTheta);
    return r;
}

// ---------------------------------------------------------------------------
// Emitter sampling — include/rt_emitter_sampling.cuh
// ---------------------------------------------------------------------------
static inline int sampleCdfReuse(const float* cdf, int count, float& u, float& outPdf) {  // :33-65
    const float sum = cdf[count];
    if (sum <= 0.0f || count <= 0) { outPdf = 0.0f; u = 0.0f; return 0; }
    const float x = u * sum;
    int lo = 0, hi = count;
    while (lo + 1 < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] <= x) lo = mid; else hi = mid;
    }
    const int idx = lo;
    const float c0 = cdf[idx], c1 = cdf[idx + 1];
    const float denom = c1 - c0;
    if (denom <= 0.0f) { u = 0.0f; outPdf = 0.0f; return idx; }
#if ORA_PLANT != 4
    u = (x - c0) / denom;                               // (planted 4: u0 not rescaled)
#endif
    outPdf = denom / sum;
    return idx;
}
struct EmitterQuery { F3 originPoint{0,0,0}, hitPoint{0,0,0}, hitNormal{0,0,0}; float pdf = 0, emitterPdf = 0; F3 dirToLight{0,0,0}; };

static inline F3 checkRadiance(const F3& radiance, const EmitterQuery& q) {  // :112-125
    const F3 negDir = {-q.dirToLight.x, -q.dirToLight.y, -q.dirToLight.z};
    if (dot3(q.hitNormal, negDir) <= 0.0f) return {0, 0, 0};
    return radiance;
}
// sampleGenerator + sampleRandomEmitter + sampleMeshEmitter (:67-100, :175-224)
static inline F3 sampleGenerator(const Scene& s, EmitterQuery& q, float u0, float u1) {
    q.emitterPdf = 0.0f;
    if (s.d.emitterCount <= 0) return {0, 0, 0};
    float epdf = 0.0f;
    const int eIdx = sampleCdfReuse(s.d.sceneEmitterCdf, s.d.emitterCount, u0, epdf);
    const OraEmitter& e = s.d.emitters[eIdx];
    // mesh sample
    F3 pos = {0, 0, 0}, nrm = {0, 0, 0};
    float meshPdf = 0.0f;
    if (e.triCount > 0) {
        const float* triCdf = s.d.emitterTriCdf + e.cdfOffset;
        float triPickPdf = 0.0f;
        const int localIdx = sampleCdfReuse(triCdf, e.triCount, u0, triPickPdf);
        const int triIdx = s.d.emitterTriIndices[e.triIndexOffset + localIdx];
        const OraTriangle& tri = s.d.triangles[triIdx];
        const float rootTerm = std::sqrt(std::fmax(0.0f, 1.0f - u0));
        const float alpha = 1.0f - rootTerm;
        const float beta = u1 * rootTerm;
        const float gamma = 1.0f - alpha - beta;
        pos = add3(add3(mul3(tri.v0, alpha), mul3(tri.v1, beta)), mul3(tri.v2, gamma));
        nrm = normalize3(cross3(sub3(tri.v1, tri.v0), sub3(tri.v2, tri.v0)));
        const float totalArea = triCdf[e.triCount];
        if (totalArea > 0.0f) meshPdf = 1.0f / totalArea;
    }
    q.hitPoint = pos;
    q.hitNormal = nrm;
    q.pdf = meshPdf;
    q.dirToLight = normalize3(sub3(q.hitPoint, q.originPoint));
    q.emitterPdf = epdf;
    return checkRadiance(e.radiance, q);
}

// ---------------------------------------------------------------------------
// Direct lighting — include/rt_direct_lighting.cuh
// ---------------------------------------------------------------------------
struct PathState {
    F3 o, d, throughput, accum;
    int bounce;
    float eta, brdfPDF;
    bool specular;
};
struct Its { F3 p, n; int tri; };
struct Shadow { F3 o, d; float tMax; F3 contrib; };

static inline float areaToSolidAngle(float pdfArea, const EmitterQuery& q) {  // :36-49
    const F3 diff = sub3(q.hitPoint, q.originPoint);
    const float dist2 = dot3(diff, diff);
    const F3 negL = {-q.dirToLight.x, -q.dirToLight.y, -q.dirToLight.z};
    const float denom = std::fmax(std::fabs(dot3(q.hitNormal, negL)), 1e-3f);
    return pdfArea * dist2 / denom;
}
static inline void accumulateEmitterHit(const Scene& s, PathState& ps, const Its& its) {  // :60-99
    const F3 surfaceRadiance = s.d.triangleEmission ? s.d.triangleEmission[its.tri] : F3{0, 0, 0};
    EmitterQuery q;
    q.originPoint = ps.o;
    q.hitPoint = its.p;
    q.hitNormal = its.n;
    q.dirToLight = ps.d;
    float brdfWeight = 1.0f;
    if (!ps.specular) {
        const int eIdx = s.emitterOfTri.empty() ? -1 : s.emitterOfTri[its.tri];
        if (eIdx >= 0) {
            const OraEmitter& e = s.d.emitters[eIdx];
            const float areaPdf = (e.areaSum <= 0.0f) ? 0.0f : 1.0f / e.areaSum;
            float selPdf = 0.0f;
            {
                const float sum = s.d.sceneEmitterCdf[s.d.emitterCount];
                if (sum > 0.0f) selPdf = (s.d.sceneEmitterCdf[eIdx + 1] - s.d.sceneEmitterCdf[eIdx]) / sum;
            }
            const float lightPdfArea = areaPdf * selPdf;
            const float lightPdfSA = areaToSolidAngle(lightPdfArea, q);
            const float denom = ps.brdfPDF + lightPdfSA;
            brdfWeight = (denom <= 0.0f) ? 0.0f : ps.brdfPDF / denom;
        }
    }
    const F3 rad = checkRadiance(surfaceRadiance, q);
    ps.accum = add3(ps.accum, mul3(mul3(rad, ps.throughput), brdfWeight));
}
static inline bool prepareAreaNEE(const Scene& s, const PathState& ps, Rng& rng, const Its& its,
                                  const OraBsdf& bsdf, Shadow& sh) {  // :107-164
    const float uL0 = rngNextFloat01(rng);
    const float uL1 = rngNextFloat01(rng);
    EmitterQuery q;
    q.originPoint = its.p;
    const F3 le = sampleGenerator(s, q, uL0, uL1);
    if ((le.x <= 0.0f && le.y <= 0.0f && le.z <= 0.0f) || q.emitterPdf <= 0.0f || q.pdf <= 0.0f) return false;
    const float distToLight = distance3(q.hitPoint, its.p);
    const float geo = std::fabs(dot3(its.n, q.dirToLight)) *
                      std::fabs(dot3(q.hitNormal, mul3(q.dirToLight, -1.0f))) / (distToLight * distToLight);
    BRec br;
    br.wi = toLocalFromNormal(its.n, mul3(ps.d, -1.0f));
    br.wo = toLocalFromNormal(its.n, q.dirToLight);
    br.measure = BSDF_ESolidAngle;
    const F3 fr = bsdfEval(bsdf, br);
    const float lightPdfArea = q.emitterPdf * q.pdf;
    const float lightPdfSA = areaToSolidAngle(lightPdfArea, q);
    const float brdfPdf = bsdfPdf(bsdf, br);
    const float denom = brdfPdf + lightPdfSA;
    if (denom <= 0.0f) return false;
    const float lightWeight = lightPdfSA / denom;
    F3 c = mul3(fr, le);
    c = mul3(c, lightWeight);
    c = mul3(c, geo);
    c = div3(c, lightPdfArea);
    sh.o = its.p;
    sh.d = q.dirToLight;
    sh.tMax = distToLight - RT_EPSILON;
    sh.contrib = mul3(c, ps.throughput);
    return true;
}
static inline bool prepareSpotNEE(const Scene& s, const PathState& ps, const Its& its, const OraBsdf& bsdf,
                                  int si, Shadow& sh) {  // :175-230
    const OraSpotlight& spot = s.d.spotlights[si];
    const F3 originToLight = sub3(spot.position, its.p);
    const float sqDist = dot3(originToLight, originToLight);
    if (sqDist < RT_EPSILON * RT_EPSILON) return false;
    const float dist = std::sqrt(sqDist);
    const F3 dirToLight = div3(originToLight, dist);
    const float lightCos = dot3(mul3(dirToLight, -1.0f), spot.direction);
    F3 spotLe;
    if (lightCos <= spot.outerConeCosine) return false;
    if (lightCos >= spot.innerConeCosine) {
        spotLe = mul3(spot.radiance, spot.intensity / sqDist);
    } else {
        const float falloff = (lightCos - spot.outerConeCosine) / (spot.innerConeCosine - spot.outerConeCosine);
#if ORA_PLANT == 5
        const float sm = falloff;                       // planted: linear falloff
#else
        const float sm = (falloff * falloff) * (3.0f - (2.0f * falloff));
#endif
        spotLe = mul3(spot.radiance, (spot.intensity * sm) / sqDist);
    }
    BRec br;
    br.wi = toLocalFromNormal(its.n, mul3(ps.d, -1.0f));
    br.wo = toLocalFromNormal(its.n, dirToLight);
    br.measure = BSDF_ESolidAngle;
    const F3 fr = bsdfEval(bsdf, br);
    const float cosSurface = std::fabs(dot3(its.n, dirToLight));
    sh.o = its.p;
    sh.d = dirToLight;
    sh.tMax = dist - RT_EPSILON;
    sh.contrib = mul3(mul3(mul3(fr, spotLe), cosSurface), ps.throughput);
    return true;
}
static inline void accumulateEnvHit(const Scene& s, PathState& ps) {  // :237-259
    if (!s.hasEnv) return;
    const F3 Le = envEval(s, ps.d);
    float w = 1.0f;
    if (!ps.specular) {
        const float ep = envPdf(s, ps.d);
        const float sum = ps.brdfPDF + ep;
        w = (sum > 0.0f) ? ps.brdfPDF / sum : 0.0f;
    }
    ps.accum = add3(ps.accum, mul3(mul3(Le, ps.throughput), w));
}
static inline bool prepareEnvNEE(const Scene& s, const PathState& ps, Rng& rng, const Its& its,
                                 const OraBsdf& bsdf, Shadow& sh) {  // :267-315
    if (!s.hasEnv) return false;
    const float u0 = rngNextFloat01(rng);
    const float u1 = rngNextFloat01(rng);
    const EnvSample es = envSample(s, u0, u1);
    if (es.pdf <= 0.0f) return false;
    if (es.radiance.x <= 0.0f && es.radiance.y <= 0.0f && es.radiance.z <= 0.0f) return false;
    const float cosAt = dot3(its.n, es.direction);
    if (cosAt <= 0.0f) return false;
    BRec br;
    br.wi = toLocalFromNormal(its.n, mul3(ps.d, -1.0f));
    br.wo = toLocalFromNormal(its.n, es.direction);
    br.measure = BSDF_ESolidAngle;
    const F3 fr = bsdfEval(bsdf, br);
    const float bp = bsdfPdf(bsdf, br);
    const float sum = bp + es.pdf;
    if (sum <= 0.0f) return false;
    const float lightWeight = es.pdf / sum;
    F3 c = mul3(fr, es.radiance);
    c = mul3(c, cosAt);
    c = div3(c, es.pdf);
    c = mul3(c, lightWeight);
    sh.o = its.p;
    sh.d = es.direction;
    sh.tMax = 1e30f - RT_EPSILON;
    sh.contrib = mul3(c, ps.throughput);
    return true;
}

// ---------------------------------------------------------------------------
// Integrator — rt_path_integrator.cuh, rt_optix_shading.cuh:213-248,
// src/optix_programs.cu:54-108 (one path per (pixel, frameIndex))
// ---------------------------------------------------------------------------
static inline void scatterPath(PathState& ps, Rng& rng, const Its& its, const OraBsdf& bsdf) {  // :25-47
    BRec br;
    br.wi = toLocalFromNormal(its.n, mul3(ps.d, -1.0f));
    br.rndExtra = rngNextFloat01(rng);
    const float u0 = rngNextFloat01(rng);
    const float u1 = rngNextFloat01(rng);
    const F3 w = bsdfSample(bsdf, br, u0, u1);
    ps.brdfPDF = bsdfPdf(bsdf, br);
    const F3 bounceDir = normalize3(toWorldFromNormal(its.n, br.wo));
    ps.o = add3(its.p, mul3(bounceDir, RT_EPSILON));
    ps.d = bounceDir;
    ps.throughput = mul3(ps.throughput, w);
    ps.eta *= br.eta;
}
static inline bool russianRoulette(PathState& ps, Rng& rng) {       // :54-65
    if (ps.bounce > 3) {
        const float q = std::fmin(maxCoeff3(ps.throughput) * ps.eta * ps.eta, 0.99f);
        if (rngNextFloat01(rng) >= q) return false;
        ps.throughput = div3(ps.throughput, q);
    }
    return true;
}

struct SampleTrace {   // optional per-bounce record for parity debugging
    int maxRec;
    int* tris;         // [maxRec] hit triangle per bounce (-1 miss, -2 not reached)
    float* aov;        // optional [6]: albedo xyz and shading normal xyz at the first
                       // non-delta hit (zeros when the path ends before one)
};

// RGBA8 texture lookup (the texture unit's bilinear filter with wrap
// addressing and normalized-float reads, src/host/scene_upload.cu:250-259),
// evaluated in exact binary32: texel centres at (i + 0.5) / size, bytes / 255.
static inline void texBilinear(const uint32_t* texels, const int* info, float u, float v, float out[4]) {
    const int W = info[1], H = info[2];
    u = u - std::floor(u);
    v = v - std::floor(v);
    if (!(u == u)) u = 0.0f;
    if (!(v == v)) v = 0.0f;
    const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = x - fx, b = y - fy;
    int ix0 = (int)fx, iy0 = (int)fy;
    if (ix0 < 0) ix0 += W;
    if (iy0 < 0) iy0 += H;
    if (ix0 >= W) ix0 -= W;
    if (iy0 >= H) iy0 -= H;
    const int ix1 = ix0 + 1 == W ? 0 : ix0 + 1, iy1 = iy0 + 1 == H ? 0 : iy0 + 1;
    const uint32_t* t = texels + info[0];
    const uint32_t t00 = t[(size_t)iy0 * W + ix0], t10 = t[(size_t)iy0 * W + ix1];
    const uint32_t t01 = t[(size_t)iy1 * W + ix0], t11 = t[(size_t)iy1 * W + ix1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int c = 0; c < 4; ++c) {
        const float c00 = (float)((t00 >> (8 * c)) & 0xffu) / 255.0f, c10 = (float)((t10 >> (8 * c)) & 0xffu) / 255.0f;
        const float c01 = (float)((t01 >> (8 * c)) & 0xffu) / 255.0f, c11 = (float)((t11 >> (8 * c)) & 0xffu) / 255.0f;
        out[c] = c00 * w00 + c10 * w10 + c01 * w01 + c11 * w11;
    }
}

// loadBsdfForTriangle (include/rt_material.cuh:15-59): the BSDF by value, with
// the albedo map replacing p0.xyz (microfacet ks recomputed) and the
// metallic-roughness map scaling a Disney BSDF's p0.w (by G) and p1.x (by B).
static inline OraBsdf loadBsdf(const Scene& s, const Hit& h) {
    OraBsdf b = s.d.bsdfs[s.d.triangleBsdfIds[h.tri]];
    if (!s.d.triangleMaterialIds || s.d.texCount <= 0) return b;
    const int m = s.d.triangleMaterialIds[h.tri];
    if (m < 0 || m >= s.d.texCount) return b;
    const OraTriangle& tri = s.d.triangles[h.tri];
    const float u = h.u, v = h.v;
    const float w0 = 1.0f - u - v;
    const float tu = w0 * tri.uv0.x + u * tri.uv1.x + v * tri.uv2.x;    // src/optix_programs.cu:137-140
    const float tv = w0 * tri.uv0.y + u * tri.uv1.y + v * tri.uv2.y;
    const int* ti = s.d.texInfo + 4 * (size_t)m;
    if (ti[1] > 0) {
        float t[4];
        texBilinear(s.d.texels, ti, tu, tv, t);
        b.p0.x = t[0];
        b.p0.y = t[1];
        b.p0.z = t[2];
        if (b.type == BSDF_Microfacet) b.p1.w = 1.0f - std::max(b.p0.x, std::max(b.p0.y, b.p0.z));
    }
    if (m < s.d.mrCount && b.type == BSDF_Disney) {
        const int* mi = s.d.mrInfo + 4 * (size_t)m;
        if (mi[1] > 0) {
            float t[4];
            texBilinear(s.d.mrTexels, mi, tu, tv, t);
            b.p0.w = b.p0.w * t[1];
            b.p1.x = b.p1.x * t[2];
        }
    }
    return b;
}

static inline Its makeIts(const Scene& s, const PathState& ps, const Hit& h) {
    Its its;
    its.tri = h.tri;
    its.p = {ps.o.x + h.t * ps.d.x, ps.o.y + h.t * ps.d.y, ps.o.z + h.t * ps.d.z};
    const OraTriangle& tri = s.d.triangles[h.tri];
    const float u = h.u, v = h.v;
    const float w0 = 1.0f - u - v;
    its.n = normalize3({w0 * tri.n0.x + u * tri.n1.x + v * tri.n2.x,
                        w0 * tri.n0.y + u * tri.n1.y + v * tri.n2.y,
                        w0 * tri.n0.z + u * tri.n1.z + v * tri.n2.z});
    return its;
}

static inline F3 primaryRayDir(const OraCamera& cam, int px, int py, int W, int H, float jx, float jy) {  // rt_camera.cuh:15-37
    const float fx = ((float)px + 0.5f + jx) / (float)W;
    const float fy = ((float)py + 0.5f + jy) / (float)H;
    const float ndcX = 2.0f * fx - 1.0f;
    const float ndcY = 2.0f * fy - 1.0f;
    const float tanHalf = om_tan(0.5f * cam.fovYRadians);
    const float sx = ndcX * cam.aspect * tanHalf;
    const float sy = ndcY * tanHalf;
    const F3 dir = add3(cam.forward, add3(mul3(cam.right, sx), mul3(cam.up, sy)));
    return normalize3(dir);
}

// One path from pixel (px, py) continuing the caller's RNG stream
// (src/optix_programs.cu:69-98 loop body).
static F3 tracePath(const Scene& s, int px, int py, int W, int H, Rng& rng, int maxDepth, SampleTrace* dbg) {
    const float jx = rngNextFloat01(rng) - 0.5f;
    const float jy = rngNextFloat01(rng) - 0.5f;
    PathState ps;
    ps.o = s.d.camera.origin;
    ps.d = primaryRayDir(s.d.camera, px, py, W, H, jx, jy);
    ps.throughput = {1.0f, 1.0f, 1.0f};
    ps.accum = {0.0f, 0.0f, 0.0f};
    ps.bounce = 0;
    ps.eta = 1.0f;
    ps.brdfPDF = 0.0f;
    ps.specular = true;
    if (dbg) for (int i = 0; i < dbg->maxRec; ++i) dbg->tris[i] = -2;
    bool aovDone = !(dbg && dbg->aov);
    if (!aovDone) for (int i = 0; i < 6; ++i) dbg->aov[i] = 0.0f;
    while (ps.bounce < maxDepth) {
        const Hit h = closestHit(s, ps.o, ps.d, RT_EPSILON, 1e30f);
        if (dbg && ps.bounce < dbg->maxRec) dbg->tris[ps.bounce] = h.tri;
        if (h.tri < 0) {
            accumulateEnvHit(s, ps);
            break;
        }
        const Its its = makeIts(s, ps, h);
        const OraBsdf bsdf = loadBsdf(s, h);
        if (s.d.triangleEmitterFlags && s.d.triangleEmitterFlags[its.tri] != 0) accumulateEmitterHit(s, ps, its);
        if (bsdfIsDiffuse(bsdf)) {
            if (!aovDone) {
                const float a[6] = {bsdf.p0.x, bsdf.p0.y, bsdf.p0.z, its.n.x, its.n.y, its.n.z};
                for (int i = 0; i < 6; ++i) dbg->aov[i] = a[i];
                aovDone = true;
            }
            ps.specular = false;
            Shadow sh;
            if (prepareAreaNEE(s, ps, rng, its, bsdf, sh) && !occluded(s, sh.o, sh.d, sh.tMax))
                ps.accum = add3(ps.accum, sh.contrib);
            for (int si = 0; si < s.d.spotlightCount; ++si)
                if (prepareSpotNEE(s, ps, its, bsdf, si, sh) && !occluded(s, sh.o, sh.d, sh.tMax))
                    ps.accum = add3(ps.accum, sh.contrib);
            if (prepareEnvNEE(s, ps, rng, its, bsdf, sh) && !occluded(s, sh.o, sh.d, sh.tMax))
                ps.accum = add3(ps.accum, sh.contrib);
        } else {
            ps.specular = true;
        }
        scatterPath(ps, rng, its, bsdf);
        if (!russianRoulette(ps, rng)) break;
        ps.bounce++;
    }
    return ps.accum;
}

// Wavefront semantics: one path per (pixel, frameIndex) with its own seed.
static F3 traceSample(const Scene& s, int px, int py, int W, int H, uint32_t frameIndex, int maxDepth,
                      SampleTrace* dbg) {
    const uint32_t pixel = (uint32_t)py * (uint32_t)W + (uint32_t)px;
    Rng rng = makeRng(hashUint(pixel ^ (frameIndex * 0x9e3779b9u)));
    return tracePath(s, px, py, W, H, rng, maxDepth, dbg);
}

static inline uint32_t packPixelGamma(const F3& lin) {               // rt_path_integrator.cuh:93-105
    const float g = 1.0f / 2.2f;
    const float dr = om_pow(std::fmax(0.0f, std::fmin(lin.x, 1.0f)), g);
    const float dg = om_pow(std::fmax(0.0f, std::fmin(lin.y, 1.0f)), g);
    const float db = om_pow(std::fmax(0.0f, std::fmin(lin.z, 1.0f)), g);
    const uint32_t r = (uint32_t)(uint8_t)(dr * 255.0f);
    const uint32_t gg = (uint32_t)(uint8_t)(dg * 255.0f);
    const uint32_t b = (uint32_t)(uint8_t)(db * 255.0f);
    return (255u << 24) | (b << 16) | (gg << 8) | r;
}

}  // namespace ora

using namespace ora;

// ===========================================================================
// C ABI used by the Python tests (ctypes)
// ===========================================================================
extern "C" {

// ---- math / rng -------------------------------------------------------------
float ora_sin(float x) { return om_sin(x); }
float ora_cos(float x) { return om_cos(x); }
float ora_tan(float x) { return om_tan(x); }
float ora_exp(float x) { return om_exp(x); }
float ora_log(float x) { return om_log(x); }
float ora_pow(float x, float y) { return om_pow(x, y); }
float ora_atan2(float y, float x) { return om_atan2(y, x); }
float ora_acos(float x) { return om_acos(x); }
void ora_math_batch(int fn, const float* x, const float* y, float* out, long n) {
    for (long i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = om_sin(x[i]); break;
            case 1: out[i] = om_cos(x[i]); break;
            case 2: out[i] = om_tan(x[i]); break;
            case 3: out[i] = om_exp(x[i]); break;
            case 4: out[i] = om_log(x[i]); break;
            case 5: out[i] = om_pow(x[i], y[i]); break;
            case 6: out[i] = om_atan2(x[i], y[i]); break;
            case 7: out[i] = om_acos(x[i]); break;
            default: out[i] = 0.0f;
        }
    }
}
uint32_t ora_hash(uint32_t x) { return hashUint(x); }
// Texture lookups for tests: info = (offset, width, height, 0); out4 per (u, v).
void ora_texture_sample(const uint32_t* texels, const int* info, const float* uv, long n, float* out4) {
    for (long i = 0; i < n; ++i) texBilinear(texels, info, uv[2 * i], uv[2 * i + 1], out4 + 4 * i);
}
// Writes n successive rngNextU32 outputs from makeRng(seed).
void ora_rng_u32(uint32_t seed, uint32_t* out, int n) {
    Rng r = makeRng(seed);
    for (int i = 0; i < n; ++i) out[i] = rngNextU32(r);
}
void ora_rng_f01(uint32_t seed, float* out, int n) {
    Rng r = makeRng(seed);
    for (int i = 0; i < n; ++i) out[i] = rngNextFloat01(r);
}
uint32_t ora_pixel_seed(uint32_t pixel, uint32_t frame) { return hashUint(pixel ^ (frame * 0x9e3779b9u)); }

// ---- bsdf -------------------------------------------------------------------
void ora_bsdf_eval(const OraBsdf* b, const float* wi, const float* wo, float* out3) {
    BRec r; r.wi = {wi[0], wi[1], wi[2]}; r.wo = {wo[0], wo[1], wo[2]}; r.measure = BSDF_ESolidAngle;
    const F3 f = bsdfEval(*b, r);
    out3[0] = f.x; out3[1] = f.y; out3[2] = f.z;
}
float ora_bsdf_pdf(const OraBsdf* b, const float* wi, const float* wo) {
    BRec r; r.wi = {wi[0], wi[1], wi[2]}; r.wo = {wo[0], wo[1], wo[2]}; r.measure = BSDF_ESolidAngle;
    return bsdfPdf(*b, r);
}
// out: [0..2] weight, [3..5] wo, [6] eta, [7] measure, [8] pdf (separate bsdfPdf call, as scatterPath does)
void ora_bsdf_sample(const OraBsdf* b, const float* wi, float rndExtra, float u0, float u1, float* out) {
    BRec r; r.wi = {wi[0], wi[1], wi[2]}; r.rndExtra = rndExtra;
    const F3 w = bsdfSample(*b, r, u0, u1);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
    out[3] = r.wo.x; out[4] = r.wo.y; out[5] = r.wo.z;
    out[6] = r.eta; out[7] = (float)r.measure;
    out[8] = bsdfPdf(*b, r);
}
void ora_bsdf_batch_sample(const OraBsdf* b, const float* wi, const float* rnd, long n, float* out) {
    for (long i = 0; i < n; ++i) ora_bsdf_sample(b, wi + 3 * i, rnd[3 * i], rnd[3 * i + 1], rnd[3 * i + 2], out + 9 * i);
}
void ora_bsdf_batch_eval(const OraBsdf* b, const float* wi, const float* wo, long n, float* out4) {
    for (long i = 0; i < n; ++i) {
        ora_bsdf_eval(b, wi + 3 * i, wo + 3 * i, out4 + 4 * i);
        out4[4 * i + 3] = ora_bsdf_pdf(b, wi + 3 * i, wo + 3 * i);
    }
}
float ora_fresnel_dielectric(float c, float etaI, float etaT) { return fresnelDielectric(c, etaI, etaT); }
void ora_frame(const float* n, const float* v, float* local3, float* world3) {
    const F3 nn = {n[0], n[1], n[2]}, vv = {v[0], v[1], v[2]};
    const F3 l = toLocalFromNormal(nn, vv);
    const F3 w = toWorldFromNormal(nn, vv);
    local3[0] = l.x; local3[1] = l.y; local3[2] = l.z;
    world3[0] = w.x; world3[1] = w.y; world3[2] = w.z;
}

// ---- camera -----------------------------------------------------------------
void ora_primary_ray(const OraCamera* cam, int px, int py, int W, int H, float jx, float jy, float* dir3) {
    const F3 d = primaryRayDir(*cam, px, py, W, H, jx, jy);
    dir3[0] = d.x; dir3[1] = d.y; dir3[2] = d.z;
}

// ---- env-map CDF construction ---------------------------------------------------
// mode 0: binary32 sequential (src/host/scene_upload.cu:428-459 exactly);
// mode 1: binary64 sequential sums rounded to binary32 at each stored entry.
float ora_env_build_cdf(const float* rgba, int W, int H, int mode, float* marg, float* cond) {
    if (mode == 0) {
        const float PI = 3.14159265358979323846f;
        marg[0] = 0.0f;
        for (int iy = 0; iy < H; ++iy) {
            const float theta = PI * ((float)iy + 0.5f) / (float)H;
            const float sinTheta = (float)std::sin((double)theta);   // host libm sinf
            float* row = cond + (size_t)iy * (size_t)(W + 1);
            row[0] = 0.0f;
            for (int ix = 0; ix < W; ++ix) {
                const float* p = rgba + 4 * ((size_t)iy * W + ix);
                const float lum = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
                row[ix + 1] = row[ix] + lum * sinTheta;
            }
            marg[iy + 1] = marg[iy] + row[W];
        }
        return marg[H];
    }
    double m = 0.0;
    marg[0] = 0.0f;
    for (int iy = 0; iy < H; ++iy) {
        const double theta = M_PI * ((double)iy + 0.5) / (double)H;
        const double st = std::sin(theta);
        float* row = cond + (size_t)iy * (size_t)(W + 1);
        double acc = 0.0;
        row[0] = 0.0f;
        for (int ix = 0; ix < W; ++ix) {
            const float* p = rgba + 4 * ((size_t)iy * W + ix);
            const double lum = 0.2126 * (double)p[0] + 0.7152 * (double)p[1] + 0.0722 * (double)p[2];
            acc += lum * st;
            row[ix + 1] = (float)acc;
        }
        m += acc;
        marg[iy + 1] = (float)m;
    }
    return (float)m;
}

// ---- scene ------------------------------------------------------------------------
void* ora_scene_create(const OraSceneDesc* d) {
    Scene* s = new Scene();
    s->d = *d;
    s->hasEnv = d->envRGBA != nullptr && d->envWidth > 0 && d->envHeight > 0;
    s->envRot = d->envRotationDeg * (3.14159265358979323846f / 180.0f);
    if (d->emitterCount > 0 && d->emitters && d->emitterTriIndices) {
        s->emitterOfTri.assign((size_t)d->triangleCount, -1);
        for (int e = 0; e < d->emitterCount; ++e) {
            const OraEmitter& E = d->emitters[e];
            for (int j = 0; j < E.triCount; ++j) {
                const int t = d->emitterTriIndices[E.triIndexOffset + j];
                if (t >= 0 && t < d->triangleCount && s->emitterOfTri[t] < 0) s->emitterOfTri[t] = e;
            }
        }
    }
    buildBvh(*s);
    return s;
}
void ora_scene_destroy(void* p) { delete static_cast<Scene*>(p); }

// Closest hit for n rays (rays: 6 floats o,d).  brute != 0 uses the O(T) loop.
void ora_trace_closest(void* p, const float* rays, long n, float tmin, float tmax, int brute,
                       int* tri, float* t, float* u, float* v) {
    const Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        const F3 o = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
        const F3 d = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
        const Hit h = brute ? bruteClosest(s, o, d, tmin, tmax) : closestHit(s, o, d, tmin, tmax);
        tri[i] = h.tri; t[i] = h.t; u[i] = h.u; v[i] = h.v;
    }
}
void ora_trace_occluded(void* p, const float* rays, const float* tmax, long n, uint8_t* occ) {
    const Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        const F3 o = {rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
        const F3 d = {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
        occ[i] = occluded(s, o, d, tmax[i]) ? 1 : 0;
    }
}

// Env helpers for tests.
void ora_env_eval(void* p, const float* dir, float* out3) {
    const F3 r = envEval(*static_cast<Scene*>(p), {dir[0], dir[1], dir[2]});
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
}
float ora_env_pdf(void* p, const float* dir) { return envPdf(*static_cast<Scene*>(p), {dir[0], dir[1], dir[2]}); }
void ora_env_sample(void* p, float u0, float u1, float* out7) {
    const EnvSample e = envSample(*static_cast<Scene*>(p), u0, u1);
    out7[0] = e.direction.x; out7[1] = e.direction.y; out7[2] = e.direction.z; out7[3] = e.pdf;
    out7[4] = e.radiance.x; out7[5] = e.radiance.y; out7[6] = e.radiance.z;
}
void ora_dir_to_equirect(const float* dir, float rotRad, float* uv) {
    const F2 r = dirToEquirect({dir[0], dir[1], dir[2]}, rotRad);
    uv[0] = r.x; uv[1] = r.y;
}
// Emitter NEE sample at origin: out = [pos3, normal3, pdf, emitterPdf, le3]
void ora_sample_emitter(void* p, const float* origin, float u0, float u1, float* out11) {
    const Scene& s = *static_cast<Scene*>(p);
    EmitterQuery q; q.originPoint = {origin[0], origin[1], origin[2]};
    const F3 le = sampleGenerator(s, q, u0, u1);
    out11[0] = q.hitPoint.x; out11[1] = q.hitPoint.y; out11[2] = q.hitPoint.z;
    out11[3] = q.hitNormal.x; out11[4] = q.hitNormal.y; out11[5] = q.hitNormal.z;
    out11[6] = q.pdf; out11[7] = q.emitterPdf; out11[8] = le.x; out11[9] = le.y; out11[10] = le.z;
}

// Spotlight NEE record (prepareSpotlightNEE) at point p3 with shading normal
// n3 for a path arriving along dIn3 with throughput thr3:
// out8 = [ok, direction3, tMax, contribution3].
void ora_spot_nee(void* p, const float* p3, const float* n3, const float* dIn3, const float* thr3,
                  const OraBsdf* bsdf, int si, float* out8) {
    const Scene& s = *static_cast<Scene*>(p);
    PathState ps{};
    ps.d = {dIn3[0], dIn3[1], dIn3[2]};
    ps.throughput = {thr3[0], thr3[1], thr3[2]};
    Its its{};
    its.p = {p3[0], p3[1], p3[2]};
    its.n = {n3[0], n3[1], n3[2]};
    Shadow sh{};
    const bool ok = si >= 0 && si < s.d.spotlightCount && prepareSpotNEE(s, ps, its, *bsdf, si, sh);
    out8[0] = ok ? 1.0f : 0.0f;
    out8[1] = sh.d.x; out8[2] = sh.d.y; out8[3] = sh.d.z; out8[4] = sh.tMax;
    out8[5] = sh.contrib.x; out8[6] = sh.contrib.y; out8[7] = sh.contrib.z;
}

// Render frames [frame0, frame0+nframes) into accum (W*H*3, binary32 Welford
// running mean as rt_path_integrator.cuh:72-87; the Welford count of frame f is
// f+1, as the reference resets its frame index together with the buffer).
// Optional sumD (W*H*3 binary64) accumulates plain sums; optional rgba8 receives
// the gamma-packed display pixels after the last frame.
void ora_render(void* p, int W, int H, int frame0, int nframes, int maxDepth, float* accum, double* sumD,
                uint32_t* rgba8, int threads) {
    const Scene& s = *static_cast<Scene*>(p);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const int tilesX = (W + 15) / 16, tilesY = (H + 15) / 16;
#pragma omp parallel for schedule(dynamic, 1)
    for (int tile = 0; tile < tilesX * tilesY; ++tile) {
        const int tx = tile % tilesX, ty = tile / tilesX;
        for (int py = ty * 16; py < std::min(H, ty * 16 + 16); ++py) {
            for (int px = tx * 16; px < std::min(W, tx * 16 + 16); ++px) {
                const size_t pix = (size_t)py * W + px;
                F3 mean = {accum[3 * pix], accum[3 * pix + 1], accum[3 * pix + 2]};
                for (int f = frame0; f < frame0 + nframes; ++f) {
                    const F3 x = traceSample(s, px, py, W, H, (uint32_t)f, maxDepth, nullptr);
                    const float nn = (float)((uint32_t)f + 1u);
                    mean = {mean.x + (x.x - mean.x) / nn, mean.y + (x.y - mean.y) / nn, mean.z + (x.z - mean.z) / nn};
                    if (sumD) { sumD[3 * pix] += x.x; sumD[3 * pix + 1] += x.y; sumD[3 * pix + 2] += x.z; }
                }
                accum[3 * pix] = mean.x; accum[3 * pix + 1] = mean.y; accum[3 * pix + 2] = mean.z;
                if (rgba8) rgba8[pix] = packPixelGamma(mean);
            }
        }
    }
}

// AOV running means (albedo, shading normal of the first non-delta hit) of
// frames [frame0, frame0+nframes), with the Welford update of ora_render.
void ora_render_aov(void* p, int W, int H, int frame0, int nframes, int maxDepth, float* alb3, float* nrm3) {
    const Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            const size_t pix = (size_t)py * W + px;
            float m[6] = {alb3[3 * pix], alb3[3 * pix + 1], alb3[3 * pix + 2],
                          nrm3[3 * pix], nrm3[3 * pix + 1], nrm3[3 * pix + 2]};
            for (int f = frame0; f < frame0 + nframes; ++f) {
                float a[6];
                SampleTrace tr{0, nullptr, a};
                (void)traceSample(s, px, py, W, H, (uint32_t)f, maxDepth, &tr);
                const float nn = (float)((uint32_t)f + 1u);
                for (int i = 0; i < 6; ++i) m[i] = m[i] + (a[i] - m[i]) / nn;
            }
            for (int i = 0; i < 3; ++i) {
                alb3[3 * pix + i] = m[i];
                nrm3[3 * pix + i] = m[3 + i];
            }
        }
    }
}

// One sample of selected pixels: radiance (3 floats each), hit triangles of
// the first maxRec bounces (-1 miss, -2 not reached) and the 6 AOV floats
// (optional outputs).
void ora_trace_samples(void* p, const int* pixels, long n, int W, int H, int frame, int maxDepth, float* rad3,
                       int* tris, int maxRec, float* aov6) {
    const Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < n; ++i) {
        const int px = pixels[i] % W, py = pixels[i] / W;
        SampleTrace tr{tris ? maxRec : 0, tris ? tris + i * maxRec : nullptr, aov6 ? aov6 + 6 * i : nullptr};
        const F3 r = traceSample(s, px, py, W, H, (uint32_t)frame, maxDepth, (tris || aov6) ? &tr : nullptr);
        rad3[3 * i] = r.x; rad3[3 * i + 1] = r.y; rad3[3 * i + 2] = r.z;
    }
}

// Megakernel frames (src/optix_programs.cu:54-108): spp paths per pixel and
// frame sharing one RNG stream seeded from (pixel, frameIndex), averaged with
// div3, then the Welford update of frame f uses count f+1.
void ora_render_megakernel(void* p, int W, int H, int frame0, int nframes, int spp, int maxDepth, float* accum) {
    const Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py) {
        for (int px = 0; px < W; ++px) {
            const size_t pix = (size_t)py * W + px;
            F3 mean = {accum[3 * pix], accum[3 * pix + 1], accum[3 * pix + 2]};
            for (int f = frame0; f < frame0 + nframes; ++f) {
                Rng rng = makeRng(hashUint((uint32_t)pix ^ ((uint32_t)f * 0x9e3779b9u)));
                F3 fin = {0.0f, 0.0f, 0.0f};
                for (int i = 0; i < spp; ++i) fin = add3(fin, tracePath(s, px, py, W, H, rng, maxDepth, nullptr));
                fin = div3(fin, (float)spp);
                const float nn = (float)((uint32_t)f + 1u);
                mean = {mean.x + (fin.x - mean.x) / nn, mean.y + (fin.y - mean.y) / nn, mean.z + (fin.z - mean.z) / nn};
            }
            accum[3 * pix] = mean.x; accum[3 * pix + 1] = mean.y; accum[3 * pix + 2] = mean.z;
        }
    }
}

uint32_t ora_pack_gamma(float r, float g, float b) { return packPixelGamma({r, g, b}); }
int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
