// bte_device.h — device-side shading library of the wavefront path tracer.
//
// Restates the reference's trace-free shading headers (include/rt_*.cuh) for
// CDNA4.  Arithmetic keeps the reference's evaluation order, every library
// call is replaced by an exactly specified binary32 sequence (the build uses
// -ffp-contract=off and IEEE division/sqrt), so a CPU implementation of the
// same sequences reproduces these results bit-for-bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "bte_types.h"

namespace bte {

#define BTE_DEV __device__ __forceinline__

// ---------------------------------------------------------------------------
// Elementary functions.  Methods (published):
//   sincos: Cody-Waite reduction by pi/2 (three-part constant, first two parts
//           12 significant bits), Taylor series sin to r^9, cos to r^10.
//   exp:    x = k ln2 + r (12-bit ln2 head + tail), Taylor e^r to r^7, 2^k in
//           two exact power-of-two factors.
//   log:    x = m 2^e, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh((m-1)/(m+1)).
//   atan2:  octant reduction, atan t = pi/4 + atan((t-1)/(t+1)) for t > tan(pi/8),
//           Taylor series to t^21.   acos x = atan2(sqrt((1-x)(1+x)), x).
// ---------------------------------------------------------------------------
namespace m {
constexpr float PIO2_1 = 1.5703125f;
constexpr float PIO2_2 = 4.837512969970703e-4f;
constexpr float PIO2_3 = 7.549790126404332e-8f;
constexpr float TWO_OVER_PI = 0.6366197466850281f;
constexpr float LN2_HI = 0.693115234375f;
constexpr float LN2_LO = 3.194618329871446e-05f;
constexpr float LOG2E = 1.4426950216293335f;
constexpr float PI = 3.1415927410125732f;
constexpr float PIO2 = 1.5707963705062866f;
constexpr float PIO4 = 0.7853981852531433f;
constexpr float TAN_PI_8 = 0.4142135679721832f;
constexpr float SQRT2 = 1.41421353816986083984f;
constexpr float IF3 = 0.1666666716337204f;
constexpr float IF4 = 0.0416666679084301f;
constexpr float IF5 = 0.008333333767950535f;
constexpr float IF6 = 0.0013888889225199819f;
constexpr float IF7 = 0.00019841270113829523f;
constexpr float IF8 = 2.4801587642286904e-05f;
constexpr float IF9 = 2.7557318844628753e-06f;
constexpr float IF10 = 2.755731998149713e-07f;
}  // namespace m

BTE_DEV float u2f(uint32_t u) { return __uint_as_float(u); }
BTE_DEV uint32_t f2u(float f) { return __float_as_uint(f); }
BTE_DEV int round_to_int(float y) { return (int)(y + (y >= 0.0f ? 0.5f : -0.5f)); }

BTE_DEV void dsincos(float x, float* sp, float* cp) {
    if (!(fabsf(x) <= 3.4028234663852886e38f)) {
        *sp = u2f(0x7fc00000u);
        *cp = u2f(0x7fc00000u);
        return;
    }
    float xr = x;
    if (fabsf(xr) > 4096.0f) xr = fmodf(xr, 4.0f * m::PIO2 * 256.0f);
    const int k = round_to_int(xr * m::TWO_OVER_PI);
    const float kf = (float)k;
    const float r = ((xr - kf * m::PIO2_1) - kf * m::PIO2_2) - kf * m::PIO2_3;
    const float r2 = r * r;
    const float s = r + r * (r2 * (-m::IF3 + r2 * (m::IF5 + r2 * (-m::IF7 + r2 * m::IF9))));
    const float c = (1.0f - 0.5f * r2) + (r2 * r2) * (m::IF4 + r2 * (-m::IF6 + r2 * (m::IF8 + r2 * (-m::IF10))));
    const int q = k & 3;
    const float ss = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
    const float cc = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
    *sp = ss;
    *cp = cc;
}
BTE_DEV float dsin(float x) { float s, c; dsincos(x, &s, &c); return s; }
BTE_DEV float dtan(float x) { float s, c; dsincos(x, &s, &c); return s / c; }

BTE_DEV float pow2i(int n) { return u2f((uint32_t)(n + 127) << 23); }
BTE_DEV float dexp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return u2f(0x7f800000u);
    if (x < -103.97208404541015625f) return 0.0f;
    const int k = round_to_int(x * m::LOG2E);
    const float kf = (float)k;
    const float r = (x - kf * m::LN2_HI) - kf * m::LN2_LO;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (m::IF3 + r * (m::IF4 + r * (m::IF5 + r * (m::IF6 + r * m::IF7))))));
    const int k1 = k / 2;
    const int k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}
BTE_DEV float dlog(float x) {
    if (x != x) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return u2f(0xff800000u);
    if (x == u2f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    const uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float mm = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (mm > m::SQRT2) { mm = mm * 0.5f; e += 1; }
    const float f = mm - 1.0f;
    const float s = f / (2.0f + f);
    const float s2 = s * s;
    const float R = s2 * (0.3333333432674408f + s2 * (0.20000000298023224f + s2 * (0.1428571492433548f +
                    s2 * (0.1111111119389534f + s2 * 0.09090909361839294f))));
    const float twos = 2.0f * s;
    const float logm = twos + twos * R;
    const float ef = (float)e;
    return ef * m::LN2_HI + (logm + ef * m::LN2_LO);
}
BTE_DEV float dpow(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : u2f(0x7f800000u);
    if (x == 1.0f) return 1.0f;
    return dexp(y * dlog(x));
}
BTE_DEV float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}
BTE_DEV float atan_poly(float t) {
    const float t2 = t * t;
    const float p = -0.3333333432674408f + t2 * (0.20000000298023224f + t2 * (-0.1428571492433548f +
                    t2 * (0.1111111119389534f + t2 * (-0.09090909361839294f + t2 * (0.07692307978868484f +
                    t2 * (-0.06666667014360428f + t2 * (0.05882352963089943f + t2 * (-0.05263157933950424f +
                    t2 * 0.0476190485060215f))))))));
    return t + t * (t2 * p);
}
BTE_DEV float datan2(float y, float x) {
    if (x != x || y != y) return u2f(0x7fc00000u);
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) {
        const float base = signbit(x) ? m::PI : 0.0f;
        return signbit(y) ? -base : base;
    }
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    float a;
    if (mx == u2f(0x7f800000u)) {
        a = (mn == mx) ? m::PIO4 : 0.0f;
    } else {
        const float t = mn / mx;
        if (t > m::TAN_PI_8) a = m::PIO4 + atan_poly((t - 1.0f) / (t + 1.0f));
        else a = atan_poly(t);
    }
    if (ay > ax) a = m::PIO2 - a;
    if (x < 0.0f) a = m::PI - a;
    if (y < 0.0f) a = -a;
    return a;
}
BTE_DEV float dacos(float x) {
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return u2f(0x7fc00000u);
    const float s = sqrtf((1.0f - x) * (1.0f + x));
    return datan2(s, x);
}

// ---------------------------------------------------------------------------
// Vectors — include/rt_cuda_math.cuh
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
BTE_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
BTE_DEV V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
BTE_DEV V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
BTE_DEV V3 mul3(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
BTE_DEV V3 mul3(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
BTE_DEV V3 div3(V3 a, float s) {
    if (s == 0.0f) return {0.0f, 0.0f, 0.0f};
    const float inv = 1.0f / s;
    return {a.x * inv, a.y * inv, a.z * inv};
}
BTE_DEV float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BTE_DEV V3 cross3(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
BTE_DEV V3 normalize3(V3 v) {
    const float len2 = dot3(v, v);
    if (len2 <= 0.0f) return {0.0f, 0.0f, 0.0f};
    const float invLen = 1.0f / sqrtf(len2);
    return mul3(v, invLen);
}
BTE_DEV float length3(V3 v) { return sqrtf(fmaxf(0.0f, dot3(v, v))); }
BTE_DEV float maxCoeff3(V3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
BTE_DEV V3 neg3(V3 v) { return mul3(v, -1.0f); }
BTE_DEV V3 xyz(float4 a) { return {a.x, a.y, a.z}; }

struct Frame { V3 t, b, n; };
BTE_DEV Frame makeFrame(V3 normal) {
    Frame f;
    f.n = normalize3(normal);
    const V3 a = (fabsf(f.n.z) < 0.999f) ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    f.t = normalize3(cross3(a, f.n));
    f.b = cross3(f.n, f.t);
    return f;
}
BTE_DEV V3 toLocal(const Frame& f, V3 v) { return {dot3(v, f.t), dot3(v, f.b), dot3(v, f.n)}; }
BTE_DEV V3 toWorld(const Frame& f, V3 v) { return add3(add3(mul3(f.t, v.x), mul3(f.b, v.y)), mul3(f.n, v.z)); }

// ---------------------------------------------------------------------------
// RNG — include/rt_rng.cuh:8-46
// ---------------------------------------------------------------------------
BTE_DEV uint32_t hashUint(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
BTE_DEV uint32_t seedRng(uint32_t seed) { return seed ? seed : 1u; }
BTE_DEV float rngNext(uint32_t& s) {
    uint32_t x = s;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    s = x;
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// BSDFs — include/rt_bsdf*.cuh.  Device table entry: 4 x float4 (p0..p3) with
// the BsdfType stored in the bits of p3.w (unused by every BSDF).
// ---------------------------------------------------------------------------
constexpr float RT_PI = 3.14159265358979323846f;
constexpr float RT_INV_PI = 1.0f / RT_PI;
constexpr float AZIMUTH = 1.0f / (2.0f * RT_PI);
constexpr float RT_EPS = 1e-4f;

struct Bsdf { float4 p0, p1, p2, p3; int type; };
struct BRec { V3 wi, wo; int measure; float eta; float rndExtra; };
BTE_DEV BRec makeRec() { BRec r; r.wi = {0, 0, 0}; r.wo = {0, 0, 0}; r.measure = 0; r.eta = 0.0f; r.rndExtra = 0.0f; return r; }

BTE_DEV V3 cosineHemisphere(float u1, float u2) {
    const float r = sqrtf(fmaxf(u1, 0.0f));
    const float phi = 2.0f * RT_PI * u2;
    float s, c;
    dsincos(phi, &s, &c);
    return {r * c, r * s, sqrtf(fmaxf(0.0f, 1.0f - u1))};
}

// diffuse — rt_bsdf_diffuse.cuh:21-62
BTE_DEV V3 evalDiffuse(const Bsdf& b, const BRec& r) {
    if (r.measure != 1 || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    return mul3(xyz(b.p0), RT_INV_PI);
}
BTE_DEV float pdfDiffuse(const BRec& r) {
    if (r.measure != 1 || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    return RT_INV_PI * r.wo.z;
}
BTE_DEV V3 sampleDiffuse(const Bsdf& b, BRec& r, float u1, float u2) {
    if (r.wi.z <= 0.0f) return {0, 0, 0};
    r.measure = 1;
    r.wo = cosineHemisphere(u1, u2);
    r.eta = 1.0f;
    return xyz(b.p0);
}
// mirror — rt_bsdf_mirror.cuh:24-33
BTE_DEV V3 sampleMirror(const Bsdf& b, BRec& r) {
    r.wo = {-r.wi.x, -r.wi.y, r.wi.z};
    r.eta = 1.0f;
    r.measure = 1;
    return xyz(b.p0);
}
// dielectric — rt_bsdf_dielectric.cuh:14-103
BTE_DEV float fresnelDielectric(float cosThetaI, float etaI, float etaT) {
    if (cosThetaI < 0.0f) {
        const float tmp = etaI; etaI = etaT; etaT = tmp;
        cosThetaI = -cosThetaI;
    }
    const float sinThetaT = (etaI / etaT) * sqrtf(fmaxf(0.0f, 1.0f - cosThetaI * cosThetaI));
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = sqrtf(fmaxf(0.0f, 1.0f - sinThetaT * sinThetaT));
    const float Rs = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    const float Rp = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    return (Rs * Rs + Rp * Rp) * 0.5f;
}
BTE_DEV V3 sampleDielectric(const Bsdf& b, BRec& r, float u1) {
    r.measure = 1;
    V3 result = {1.0f, 1.0f, 1.0f};
    float cosThetaI = r.wi.z;
    const float intIOR = b.p1.x, extIOR = b.p1.y;
    const float fr = fresnelDielectric(cosThetaI, extIOR, intIOR);
    if (u1 < fr) {
        r.wo = {-r.wi.x, -r.wi.y, r.wi.z};
        r.eta = 1.0f;
    } else {
        float ratio;
        float nz = 1.0f;
        if (cosThetaI > 0) {
            ratio = extIOR / intIOR;
        } else {
            ratio = intIOR / extIOR;
            nz = -1.0f;
        }
        cosThetaI = fabsf(cosThetaI);
        const float rootTerm = 1.0f - (ratio * ratio) * (1.0f - cosThetaI * cosThetaI);
        const float dia = ratio * cosThetaI - sqrtf(rootTerm);
        r.wo = {-ratio * r.wi.x + dia * 0.0f, -ratio * r.wi.y + dia * 0.0f, -ratio * r.wi.z + dia * nz};
        r.eta = 1.0f / ratio;
        const float w = ratio * ratio;
        result = {w, w, w};
    }
    return result;
}
// microfacet — rt_bsdf_microfacet.cuh:23-185
BTE_DEV float tanThetaMf(V3 w) {
    const float ct = w.z;
    const float sin2 = fmaxf(0.0f, 1.0f - ct * ct);
    const float cos2 = ct * ct;
    if (cos2 <= 0.0f) return 0.0f;
    return sqrtf(sin2 / cos2);
}
BTE_DEV float beckmannPdf(V3 mm, float alpha) {
    if (mm.z <= 0.0f) return 0.0f;
    const float tanTheta = tanThetaMf(mm);
    const float alpha2 = alpha * alpha;
    const float cosTheta = mm.z;
    const float cosTheta3 = cosTheta * cosTheta * cosTheta;
    return AZIMUTH * (2.0f * dexp(-1.0f * tanTheta * tanTheta / alpha2)) / (alpha2 * cosTheta3);
}
BTE_DEV float shadowMaskMf(float alpha, V3 wv, V3 wh) {
    const float c = dot3(wv, wh) / wv.z;
    if (c <= 0.0f) return 0.0f;
    const float bb = 1.0f / (alpha * tanThetaMf(wv));
    if (bb < 1.6f) return ((3.535f * bb) + 2.181f * bb * bb) / (1.0f + (2.276f * bb) + 2.577f * bb * bb);
    return 1.0f;
}
BTE_DEV V3 evalMicrofacet(const Bsdf& b, const BRec& r) {
    if (r.measure != 1 || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const float alpha = b.p1.x, intIOR = b.p1.y, extIOR = b.p1.z, ks = b.p1.w;
    const V3 diffuseTerm = {b.p0.x * RT_INV_PI, b.p0.y * RT_INV_PI, b.p0.z * RT_INV_PI};
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float tanValH = tanThetaMf(wh);
    const float cosValH = wh.z;
    const float num = 2.0f * dexp(-1.0f * tanValH * tanValH / (alpha * alpha));
    const float den = alpha * alpha * cosValH * cosValH * cosValH;
    const float D = AZIMUTH * num / den;
    const float F = fresnelDielectric(dot3(wh, r.wi), extIOR, intIOR);
    const float G = shadowMaskMf(alpha, r.wi, wh) * shadowMaskMf(alpha, r.wo, wh);
    const float ecf = 4.0f * r.wi.z * cosValH * r.wo.z;
    const float spec = ks * (D * F * G) / ecf;
    return {diffuseTerm.x + spec, diffuseTerm.y + spec, diffuseTerm.z + spec};
}
BTE_DEV float pdfMicrofacet(const Bsdf& b, const BRec& r) {
    if (r.measure != 1 || r.wi.z <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const float alpha = b.p1.x, ks = b.p1.w;
    const float chp = (r.wo.z <= 0.0f) ? 0.0f : RT_INV_PI * r.wo.z;
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float bp = beckmannPdf(wh, alpha);
    const float jac = 1.0f / (4.0f * dot3(wh, r.wo));
    return ks * bp * jac + (1.0f - ks) * chp;
}
BTE_DEV V3 sampleMicrofacet(const Bsdf& b, BRec& r, float u1, float u2) {
    if (r.wi.z <= 0.0f) return {0, 0, 0};
    const float ks = b.p1.w;
    float sx = u1;
    int type;
    if (sx < ks) { type = 0; sx = sx / ks; }
    else { type = 1; sx = (sx - ks) / (1.0f - ks); }
    r.measure = 1;
    r.eta = 1.0f;
    if (type == 0) {
        const float alpha = b.p1.x;
        const float phi = 2.0f * RT_PI * u2;
        const float tan2 = -alpha * alpha * dlog(fmaxf(1.0e-6f, 1.0f - sx));
        const float ct = 1.0f / sqrtf(1.0f + tan2);
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float s, c;
        dsincos(phi, &s, &c);
        const V3 n = {st * c, st * s, ct};
        r.wo = sub3(mul3(n, 2.0f * dot3(r.wi, n)), r.wi);
    } else {
        r.wo = cosineHemisphere(sx, u2);
    }
    if (r.wo.z <= 0.0f) return {0, 0, 0};
    const float p = pdfMicrofacet(b, r);
    if (p <= 0.0f) return {0, 0, 0};
    return mul3(evalMicrofacet(b, r), r.wo.z / p);
}

// Disney — rt_bsdf_disney.cuh:16-892
BTE_DEV float dFresnel(float cosTheta, float ninety) { return 1.0f + ((ninety - 1.0f) * pow5(1.0f - cosTheta)); }
BTE_DEV V3 evalDiffuseLobe(const Bsdf& b, const BRec& r) {
    const float ci = r.wi.z, co = r.wo.z;
    if (ci <= 0.0f || co <= 0.0f) return {0, 0, 0};
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float h = fabsf(dot3(wh, r.wo));
    const float rough = b.p0.w;
    const float dfN = 0.5f + (2.0f * rough * h * h);
    const float ssN = rough * h * h;
    const V3 cop = div3(xyz(b.p0), RT_PI);
    const V3 base = mul3(cop, dFresnel(ci, dfN) * dFresnel(co, dfN) * co);
    const float ssDenom = fmaxf(1e-2f, ci + co);
    const V3 sub = mul3(cop, 1.25f * (dFresnel(ci, ssN) * dFresnel(co, ssN) * ((1.0f / ssDenom) - 0.5f) + 0.5f) * co);
    const float ss = b.p2.z;
    return add3(mul3(base, 1.0f - ss), mul3(sub, ss));
}
BTE_DEV float pdfDiffuseLobe(const BRec& r) {
    if (r.measure != 1 || r.wi.z <= 0 || r.wo.z <= 0) return 0.0f;
    return RT_INV_PI * r.wo.z;
}
BTE_DEV V3 tintColor(const Bsdf& b, float amount) {
    const V3 base = xyz(b.p0);
    const float lumi = 0.3f * base.x + 0.6f * base.y + 0.1f * base.z;
    V3 tint = {1.0f, 1.0f, 1.0f};
    if (lumi > 0) tint = div3(base, lumi);
    return {(1.0f - amount) + (amount * tint.x), (1.0f - amount) + (amount * tint.y), (1.0f - amount) + (amount * tint.z)};
}
BTE_DEV V3 fresnelMetal(const Bsdf& b, float hDotWo) {
    const V3 base = xyz(b.p0);
    const V3 spec = tintColor(b, b.p1.w);
    const float metallic = b.p1.x;
    const float en = b.p3.z - 1.0f, ed = b.p3.z + 1.0f;
    const float schlick = (en * en) / (ed * ed);
    const float coef = b.p1.y * schlick * (1.0f - metallic);
    const V3 clr = {(coef * spec.x) + (base.x * metallic), (coef * spec.y) + (base.y * metallic),
                    (coef * spec.z) + (base.z * metallic)};
    const float st = pow5(1.0f - hDotWo);
    return {clr.x + ((1.0f - clr.x) * st), clr.y + ((1.0f - clr.y) * st), clr.z + ((1.0f - clr.z) * st)};
}
BTE_DEV float ggxD(V3 wh, float ax, float ay) {
    const float q = ((wh.x * wh.x) / (ax * ax)) + ((wh.y * wh.y) / (ay * ay)) + (wh.z * wh.z);
    const float den = RT_PI * ax * ay * (q * q);
    return 1.0f / den;
}
BTE_DEV float smithG1(V3 w, float ax, float ay) {
    const float axw = w.x * ax;
    const float ayw = w.y * ay;
    const float num = sqrtf(1 + (((axw * axw) + (ayw * ayw)) / fmaxf(1e-8f, (w.z * w.z)))) - 1.0f;
    const float lambda = num / 2.0f;
    return 1.0f / (1.0f + lambda);
}
BTE_DEV void disneyAlpha(const Bsdf& b, float& ax, float& ay) {
    const float aspect = sqrtf(1.0f - (0.9f * b.p2.w));
    const float r2 = b.p0.w * b.p0.w;
    ax = fmaxf(0.0001f, r2 / aspect);
    ay = fmaxf(0.0001f, r2 * aspect);
}
BTE_DEV V3 evalMetalLobe(const Bsdf& b, const BRec& r, float ax, float ay) {
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float h = fabsf(dot3(wh, r.wo));
    const V3 fm = fresnelMetal(b, h);
    const float dm = ggxD(wh, ax, ay);
    const float gm = smithG1(r.wi, ax, ay) * smithG1(r.wo, ax, ay);
    return div3(mul3(fm, dm * gm), (4 * ci));
}
BTE_DEV float pdfMetalLobe(const BRec& r, float ax, float ay) {
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float dm = ggxD(wh, ax, ay);
    const float gm = smithG1(r.wi, ax, ay);
    const float hi = fabsf(dot3(wh, r.wi));
    const float ho = fabsf(dot3(wh, r.wo));
    return (dm * gm * hi) / (ci * 4.0f * ho);
}
BTE_DEV V3 sampleVisibleNormal(V3 w, float ax, float ay, float u1, float u2) {
    const V3 vh = normalize3(V3{ax * w.x, ay * w.y, w.z});
    const float pl = (vh.x * vh.x) + (vh.y * vh.y);
    V3 o1 = {1.0f, 0.0f, 0.0f};
    if (pl > 0.0f) o1 = div3(V3{-1.0f * vh.y, vh.x, 0.0f}, sqrtf(pl));
    const V3 o2 = cross3(vh, o1);
    const float radius = sqrtf(u1);
    const float phi = 2.0f * RT_PI * u2;
    float s, c;
    dsincos(phi, &s, &c);
    const float t1 = radius * c;
    float t2 = radius * s;
    const float skew = 0.5f * (1.0f + vh.z);
    t2 = (1.0f - skew) * sqrtf(1.0f - (t1 * t1)) + (skew * t2);
    const V3 ph = add3(add3(mul3(o1, t1), mul3(o2, t2)), mul3(vh, sqrtf(fmaxf(0.0f, 1 - (t1 * t1) - (t2 * t2)))));
    return normalize3(V3{ax * ph.x, ay * ph.y, fmaxf(0.0f, ph.z)});
}
BTE_DEV float clearcoatIr2(const Bsdf& b) {
    const float g = b.p3.y;
    const float ir = ((1.0f - g) * 0.1f) + (g * 0.001f);
    return ir * ir;
}
BTE_DEV float clearcoatD(float ir2, V3 wh) {
    const float num = ir2 - 1.0f;
    const float den = RT_PI * dlog(ir2) * (1.0f + ((ir2 - 1.0f) * (wh.z * wh.z)));
    return num / den;
}
BTE_DEV V3 evalClearcoatLobe(const Bsdf& b, const BRec& r) {
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return {0, 0, 0};
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float h = fabsf(dot3(wh, r.wo));
    const float fc = 0.04f + ((1.0f - 0.04f) * pow5(1.0f - h));
    const float dc = clearcoatD(clearcoatIr2(b), wh);
    const float gc = smithG1(r.wi, 0.25f, 0.25f) * smithG1(r.wo, 0.25f, 0.25f);
    const float cc = (fc * dc * gc) / (4.0f * ci);
    return {cc, cc, cc};
}
BTE_DEV float pdfClearcoatLobe(const Bsdf& b, const BRec& r) {
    const float ci = r.wi.z;
    if (ci <= 0.0f || r.wo.z <= 0.0f) return 0.0f;
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float dc = clearcoatD(clearcoatIr2(b), wh);
    const float h = fabsf(dot3(wh, r.wo));
    return (dc * wh.z) / (4.0f * h);
}
BTE_DEV float fresnelGlass(float hi, float ho, float ehi, float eho) {
    hi = fabsf(hi); ho = fabsf(ho); ehi = fabsf(ehi); eho = fabsf(eho);
    const float rs = (hi - eho) / (hi + eho);
    const float rp = (ehi - ho) / (ehi + ho);
    return 0.5f * ((rs * rs) + (rp * rp));
}
BTE_DEV V3 glassHalf(const BRec& r, float eta, float inNout) {
    V3 wh;
    if (inNout > 0.0f) wh = normalize3(add3(r.wi, r.wo));
    else wh = mul3(normalize3(add3(r.wi, mul3(r.wo, eta))), -1.0f);
    if (wh.z < 0.0f) wh = mul3(wh, -1.0f);
    return wh;
}
BTE_DEV V3 evalGlassLobe(const Bsdf& b, const BRec& r, float ax, float ay) {
    const float ci = r.wi.z, co = r.wo.z;
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const float inNout = ci * co;
    const V3 wh = glassHalf(r, eta, inNout);
    const float hi = dot3(wh, r.wi);
    const float ho = dot3(wh, r.wo);
    const float ehi = eta * hi;
    const float eho = eta * ho;
    const float fg = fresnelGlass(hi, ho, ehi, eho);
    const float dg = ggxD(wh, ax, ay);
    const float gg = smithG1(r.wi, ax, ay) * smithG1(r.wo, ax, ay);
    if (inNout > 0.0f) {
        return div3(mul3(xyz(b.p0), fg * dg * gg), 4.0f * fabsf(ci));
    }
    const V3 sb = {sqrtf(fmaxf(0.0f, b.p0.x)), sqrtf(fmaxf(0.0f, b.p0.y)), sqrtf(fmaxf(0.0f, b.p0.z))};
    const V3 numer = mul3(sb, (1.0f - fg) * dg * gg * fabsf(hi * ho));
    const float q = hi + eho;
    return div3(numer, fabsf(ci) * fmaxf(q * q, 1e-4f));
}
BTE_DEV float pdfGlassLobe(const Bsdf& b, const BRec& r, float ax, float ay) {
    const float ci = r.wi.z, co = r.wo.z;
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const float inNout = ci * co;
    const V3 wh = glassHalf(r, eta, inNout);
    const float hi = dot3(wh, r.wi);
    const float ho = dot3(wh, r.wo);
    const float fg = fresnelGlass(hi, ho, eta * hi, eta * ho);
    const float dg = ggxD(wh, ax, ay);
    const float gg = smithG1(r.wi, ax, ay);
    const float common = (dg * gg * fabsf(hi)) / fabsf(ci);
    if (inNout > 0.0f) return (common * fg) / (4.0f * fabsf(ho));
    const float q = hi + (eta * ho);
    const float jac = (eta * eta * fabsf(ho)) / fmaxf(q * q, 1e-4f);
    return (1.0f - fg) * common * jac;
}
BTE_DEV void sampleGlassLobe(const Bsdf& b, BRec& r, float u1, float u2, float ax, float ay) {
    const float ci = r.wi.z;
    r.measure = 1;
    float eta = b.p3.z;
    if (ci < 0.0f) eta = 1.0f / eta;
    const V3 wu = (ci < 0.0f) ? mul3(r.wi, -1.0f) : r.wi;
    V3 h = sampleVisibleNormal(wu, ax, ay, u1, u2);
    if (ci < 0.0f) h = mul3(h, -1.0f);
    const float hi = dot3(h, r.wi);
    const float s2t = (1.0f - hi * hi) / (eta * eta);
    const float ht = sqrtf(fmaxf(0.0f, 1.0f - s2t));
    const float fg = fresnelGlass(hi, ht, eta * hi, eta * ht);
    if (r.rndExtra < fg || s2t >= 1.0f) {
        r.wo = sub3(mul3(h, 2.0f * dot3(r.wi, h)), r.wi);
        r.eta = 1.0f;
    } else {
        r.wo = normalize3(sub3(mul3(h, (hi / eta) - ht), div3(r.wi, eta)));
        r.eta = eta;
    }
}
BTE_DEV V3 evalSheenLobe(const Bsdf& b, const BRec& r) {
    const float ci = r.wi.z, co = r.wo.z;
    if (ci <= 0.0f || co <= 0.0f) return {0, 0, 0};
    const V3 wh = normalize3(add3(r.wi, r.wo));
    const float h = fabsf(dot3(wh, r.wo));
    const V3 sheen = tintColor(b, b.p2.y);
    return mul3(sheen, pow5(1.0f - h) * co);
}
BTE_DEV V3 evalDisney(const Bsdf& b, const BRec& r) {
    const float specT = b.p1.z;
    float ax, ay;
    disneyAlpha(b, ax, ay);
    if (r.wi.z <= 0.0f) {
        if (specT <= 0.0f) return {0, 0, 0};
        return evalGlassLobe(b, r, ax, ay);
    }
    const float minMetal = (1.0f - b.p1.x);
    const float minSpecT = (1.0f - specT);
    // Every lobe is finite when wi.z > 0, so a lobe whose weight is exactly
    // zero contributes a signed zero to the sum; it is skipped unevaluated.
    const float wD = minMetal * minSpecT, wS = minMetal * b.p2.x, wM = (1.0f - (specT * minMetal));
    const float wC = 0.25f * b.p3.x, wG = minMetal * specT;
    const V3 z = {0.0f, 0.0f, 0.0f};
    const V3 fd = (wD != 0.0f) ? mul3(evalDiffuseLobe(b, r), wD) : z;
    const V3 fs = (wS != 0.0f) ? mul3(evalSheenLobe(b, r), wS) : z;
    const V3 fm = (wM != 0.0f) ? mul3(evalMetalLobe(b, r, ax, ay), wM) : z;
    const V3 fc = (wC != 0.0f) ? mul3(evalClearcoatLobe(b, r), wC) : z;
    const V3 fg = (wG != 0.0f) ? mul3(evalGlassLobe(b, r, ax, ay), wG) : z;
    return add3(add3(add3(add3(fd, fm), fg), fc), fs);
}
BTE_DEV float pdfDisney(const Bsdf& b, const BRec& r) {
    const float specT = b.p1.z;
    const float minMetal = (1.0f - b.p1.x);
    const float minSpecT = (1.0f - specT);
    const float dW = minMetal * minSpecT;
    const float mW = 1.0f - (specT * minMetal);
    const float gW = minMetal * specT;
    const float cW = 0.25f * b.p3.x;
    const float wSum = dW + gW + mW + cW;
    float ax, ay;
    disneyAlpha(b, ax, ay);
    if (r.wi.z <= 0) {
        if (specT <= 0.0f) return 0.0f;
        return pdfGlassLobe(b, r, ax, ay);
    }
    // zero-weight lobes skipped as in evalDisney (their pdfs are finite)
    const float pD = (dW != 0.0f) ? pdfDiffuseLobe(r) : 0.0f;
    const float pM = (mW != 0.0f) ? pdfMetalLobe(r, ax, ay) : 0.0f;
    const float pG = (gW != 0.0f) ? pdfGlassLobe(b, r, ax, ay) : 0.0f;
    const float pC = (cW != 0.0f) ? pdfClearcoatLobe(b, r) : 0.0f;
    return (dW * pD + gW * pG + mW * pM + cW * pC) / wSum;
}
// Lobe choice and lobe sampling of bsdfSampleDisney for wi.z > 0
// (rt_bsdf_disney.cuh:843-880).
BTE_DEV void sampleDisneyLobeChoice(const Bsdf& b, BRec& r, float u1, float u2, float ax, float ay) {
    const float specT = b.p1.z;
    {
        const float minMetal = (1.0f - b.p1.x);
        const float dW = minMetal * (1.0f - specT);
        const float sW = minMetal * specT;
        const float mW = 1.0f - (specT * minMetal);
        const float cW = 0.25f * b.p3.x;
        const float wSum = dW + sW + mW + cW;
        const float rc = r.rndExtra;
        const float dF = dW / wSum;
        const float sF = sW / wSum;
        const float mF = mW / wSum;
        if (rc < dF) {
            // diffuse lobe (wi.z > 0 here)
            r.measure = 1;
            r.wo = cosineHemisphere(u1, u2);
            r.eta = 1.0f;
        } else if (rc < dF + sF) {
            r.rndExtra = (sF > 0.0f) ? (rc - dF) / sF : 0.0f;
            sampleGlassLobe(b, r, u1, u2, ax, ay);
        } else if (rc < dF + sF + mF) {
            r.measure = 1;
            r.eta = 1.0f;
            const V3 h = sampleVisibleNormal(r.wi, ax, ay, u1, u2);
            r.wo = sub3(mul3(h, 2.0f * dot3(r.wi, h)), r.wi);
        } else {
            r.measure = 1;
            const float ir2 = clearcoatIr2(b);
            const float az = 2.0f * RT_PI * u1;
            const float ce2 = (1.0f - dpow(ir2, 1.0f - u2)) / (1.0f - ir2);
            const float se = sqrtf(1.0f - ce2);
            const float hz = sqrtf(ce2);
            float s, c;
            dsincos(az, &s, &c);
            const V3 h = normalize3(V3{se * c, se * s, hz});
            r.wo = sub3(mul3(h, 2.0f * dot3(r.wi, h)), r.wi);
            r.eta = 1.0f;
        }
    }
}
BTE_DEV V3 sampleDisney(const Bsdf& b, BRec& r, float u1, float u2) {
    const float specT = b.p1.z;
    float ax, ay;
    disneyAlpha(b, ax, ay);
    if (r.wi.z <= 0.0f) {
        if (specT <= 0.0f) return {0, 0, 0};
        sampleGlassLobe(b, r, u1, u2, ax, ay);
    } else {
        sampleDisneyLobeChoice(b, r, u1, u2, ax, ay);
    }
    const float pdf = pdfDisney(b, r);
    if (pdf < 1e-7f) return {0, 0, 0};
    return div3(evalDisney(b, r), pdf);
}

BTE_DEV V3 bsdfEval(const Bsdf& b, const BRec& r) {
    if (b.type == BTE_BSDF_DIFFUSE) return evalDiffuse(b, r);
    if (b.type == BTE_BSDF_MICROFACET) return evalMicrofacet(b, r);
    if (b.type == BTE_BSDF_DISNEY) return evalDisney(b, r);
    return {0, 0, 0};
}
BTE_DEV float bsdfPdf(const Bsdf& b, const BRec& r) {
    if (b.type == BTE_BSDF_DIFFUSE) return pdfDiffuse(r);
    if (b.type == BTE_BSDF_MICROFACET) return pdfMicrofacet(b, r);
    if (b.type == BTE_BSDF_DISNEY) return pdfDisney(b, r);
    return 0.0f;
}
BTE_DEV V3 bsdfSample(const Bsdf& b, BRec& r, float u1, float u2) {
    if (b.type == BTE_BSDF_DIFFUSE) return sampleDiffuse(b, r, u1, u2);
    if (b.type == BTE_BSDF_DIELECTRIC) return sampleDielectric(b, r, u1);
    if (b.type == BTE_BSDF_MIRROR) return sampleMirror(b, r);
    if (b.type == BTE_BSDF_MICROFACET) return sampleMicrofacet(b, r, u1, u2);
    if (b.type == BTE_BSDF_DISNEY) return sampleDisney(b, r, u1, u2);
    return {0, 0, 0};
}
BTE_DEV bool bsdfIsDiffuse(int type) {
    return type == BTE_BSDF_DIFFUSE || type == BTE_BSDF_MICROFACET || type == BTE_BSDF_DISNEY;
}

BTE_DEV uint32_t packGamma(V3 lin) {                               // rt_path_integrator.cuh:93-105
    const float g = 1.0f / 2.2f;
    const float dr = dpow(fmaxf(0.0f, fminf(lin.x, 1.0f)), g);
    const float dg = dpow(fmaxf(0.0f, fminf(lin.y, 1.0f)), g);
    const float db = dpow(fmaxf(0.0f, fminf(lin.z, 1.0f)), g);
    const uint32_t r = (uint32_t)(uint8_t)(dr * 255.0f);
    const uint32_t gg = (uint32_t)(uint8_t)(dg * 255.0f);
    const uint32_t b = (uint32_t)(uint8_t)(db * 255.0f);
    return (255u << 24) | (b << 16) | (gg << 8) | r;
}

}  // namespace bte
