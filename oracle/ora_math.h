// ora_math.h — ORACLE (test infrastructure only; never linked into the product).
//
// Exactly specified single-precision elementary functions for the CPU
// restatement of the reference integrator.  Every function is built from
// IEEE-754 binary32 +, -, *, /, sqrt, comparisons and float<->int conversions
// only (the build uses -ffp-contract=off), so any implementation that performs
// the same operation sequence gets identical bits.  The product's HIP kernels
// implement the same sequences independently and are checked against these
// bit-for-bit.
//
// Methods (all published, constants derived in this file's comments):
//   sin/cos/tan : Cody & Waite additive range reduction by pi/2 split into
//                 three binary32 parts (12+12+24 significant bits), then the
//                 Taylor series of sin (to r^9) and cos (to r^10) on
//                 [-pi/4, pi/4]; quadrant selection by k mod 4.
//   exp         : reduction x = k ln2 + r with ln2 split into a 12-bit head
//                 and a binary32 tail, Taylor series of e^r to r^7, scaling by
//                 2^k in two exact power-of-two steps.
//   log         : x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//                 log m = 2 atanh(s) = 2s (1 + s^2/3 + ... + s^10/11).
//   atan2       : octant reduction to t in [0,1], a second reduction
//                 atan t = pi/4 + atan((t-1)/(t+1)) for t > tan(pi/8), and the
//                 Taylor series of atan to t^21.
//   acos        : acos x = atan2(sqrt((1-x)(1+x)), x).
//   pow         : exp(y log x) (x > 0); pow5 by repeated multiplication.
//
// Accuracy against correctly rounded results is checked in
// tests/test_oracle_math.py (bounds stated there).  The reference calls the
// CUDA math library (expf, logf, powf, sinf, cosf, tanf, atan2f, acosf),
// whose exact results are not reproducible here; results that depend on the
// last bits of those functions are therefore parity-unpinned against the
// reference and pinned against this specification instead.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace ora {

static inline uint32_t f2u(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

// ---- constants ------------------------------------------------------------
// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (first two parts hold 12 significant bits so
// k*PIO2_1 and k*PIO2_2 are exact for |k| < 4096).
static constexpr float PIO2_1 = 1.5703125f;             // 0x3fc90000
static constexpr float PIO2_2 = 4.837512969970703e-4f;  // 0x39fda000
static constexpr float PIO2_3 = 7.549790126404332e-8f;  // 0x33a22169
static constexpr float TWO_OVER_PI = 0.6366197466850281f;
// ln2 = LN2_HI + LN2_LO, LN2_HI with 12 significant bits.
static constexpr float LN2_HI = 0.693115234375f;        // 0x3f317000
static constexpr float LN2_LO = 3.194618329871446e-05f; // 0x3805fdf4
static constexpr float LOG2E = 1.4426950216293335f;
static constexpr float F_PI = 3.1415927410125732f;
static constexpr float F_PIO2 = 1.5707963705062866f;
static constexpr float F_PIO4 = 0.7853981852531433f;
static constexpr float TAN_PI_8 = 0.4142135679721832f;
static constexpr float SQRT2 = 1.41421353816986083984f;
// 1/k! rounded to binary32
static constexpr float INV_F3 = 0.1666666716337204f;
static constexpr float INV_F4 = 0.0416666679084301f;
static constexpr float INV_F5 = 0.008333333767950535f;
static constexpr float INV_F6 = 0.0013888889225199819f;
static constexpr float INV_F7 = 0.00019841270113829523f;
static constexpr float INV_F8 = 2.4801587642286904e-05f;
static constexpr float INV_F9 = 2.7557318844628753e-06f;
static constexpr float INV_F10 = 2.755731998149713e-07f;

// Round-half-away-from-zero to an integer (|y| < 2^30 assumed by callers).
static inline int round_to_int(float y) {
    return (int)(y + (y >= 0.0f ? 0.5f : -0.5f));
}

// ---- sin / cos ------------------------------------------------------------
static inline void sincos_kernel(float r, float& s, float& c) {
    const float r2 = r * r;
    s = r + r * (r2 * (-INV_F3 + r2 * (INV_F5 + r2 * (-INV_F7 + r2 * INV_F9))));
    c = (1.0f - 0.5f * r2) + (r2 * r2) * (INV_F4 + r2 * (-INV_F6 + r2 * (INV_F8 + r2 * (-INV_F10))));
}

static inline void om_sincos(float x, float* sp, float* cp) {
    if (!(std::fabs(x) <= 3.4028234663852886e38f)) { // inf or nan
        const float n = u2f(0x7fc00000u);
        *sp = n; *cp = n; return;
    }
    float xr = x;
    if (std::fabs(xr) > 4096.0f) {
        // Coarse pre-reduction (exact fmod) for large arguments.
        xr = std::fmod(xr, 4.0f * F_PIO2 * 256.0f);
    }
    const int k = round_to_int(xr * TWO_OVER_PI);
    const float kf = (float)k;
    const float r = ((xr - kf * PIO2_1) - kf * PIO2_2) - kf * PIO2_3;
    float s, c;
    sincos_kernel(r, s, c);
    switch (k & 3) {
        case 0: *sp = s; *cp = c; break;
        case 1: *sp = c; *cp = -s; break;
        case 2: *sp = -s; *cp = -c; break;
        default: *sp = -c; *cp = s; break;
    }
}
static inline float om_sin(float x) { float s, c; om_sincos(x, &s, &c); return s; }
static inline float om_cos(float x) { float s, c; om_sincos(x, &s, &c); return c; }
static inline float om_tan(float x) { float s, c; om_sincos(x, &s, &c); return s / c; }

// ---- exp ------------------------------------------------------------------
static inline float pow2i(int n) { // n in [-126, 127]
    return u2f((uint32_t)(n + 127) << 23);
}
static inline float om_exp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return u2f(0x7f800000u);
    if (x < -103.97208404541015625f) return 0.0f;
    const int k = round_to_int(x * LOG2E);
    const float kf = (float)k;
    const float r = (x - kf * LN2_HI) - kf * LN2_LO;
    const float p = 1.0f + r * (1.0f + r * (0.5f + r * (INV_F3 + r * (INV_F4 + r * (INV_F5 + r * (INV_F6 + r * INV_F7))))));
    const int k1 = k / 2;
    const int k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

// ---- log ------------------------------------------------------------------
static inline float om_log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return u2f(0xff800000u);
    if (x == u2f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    const uint32_t b = f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > SQRT2) { m = m * 0.5f; e += 1; }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float s2 = s * s;
    const float R = s2 * (0.3333333432674408f + s2 * (0.20000000298023224f + s2 * (0.1428571492433548f +
                    s2 * (0.1111111119389534f + s2 * 0.09090909361839294f))));
    const float twos = 2.0f * s;
    const float logm = twos + twos * R;
    const float ef = (float)e;
    return ef * LN2_HI + (logm + ef * LN2_LO);
}

// ---- pow ------------------------------------------------------------------
static inline float om_pow(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : u2f(0x7f800000u);
    if (x == 1.0f) return 1.0f;
    return om_exp(y * om_log(x));
}
static inline float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}

// ---- atan2 / acos -----------------------------------------------------------
static inline float atan_poly(float t) { // |t| <= tan(pi/8)
    const float t2 = t * t;
    // Taylor series of atan: t - t^3/3 + t^5/5 - ... - t^19/19 + t^21/21
    const float p = -0.3333333432674408f + t2 * (0.20000000298023224f + t2 * (-0.1428571492433548f +
                    t2 * (0.1111111119389534f + t2 * (-0.09090909361839294f + t2 * (0.07692307978868484f +
                    t2 * (-0.06666667014360428f + t2 * (0.05882352963089943f + t2 * (-0.05263157933950424f +
                    t2 * 0.0476190485060215f))))))));
    return t + t * (t2 * p);
}
static inline float om_atan2(float y, float x) {
    if (x != x || y != y) return u2f(0x7fc00000u);
    const float ax = std::fabs(x);
    const float ay = std::fabs(y);
    if (ax == 0.0f && ay == 0.0f) {
        const float base = std::signbit(x) ? F_PI : 0.0f;
        return std::signbit(y) ? -base : base;
    }
    const float mx = ax > ay ? ax : ay;
    const float mn = ax > ay ? ay : ax;
    float a;
    if (mx == u2f(0x7f800000u)) {
        a = (mn == mx) ? F_PIO4 : 0.0f;
    } else {
        const float t = mn / mx;
        if (t > TAN_PI_8) {
            a = F_PIO4 + atan_poly((t - 1.0f) / (t + 1.0f));
        } else {
            a = atan_poly(t);
        }
    }
    if (ay > ax) a = F_PIO2 - a;
    if (x < 0.0f) a = F_PI - a;
    if (y < 0.0f) a = -a;
    return a;
}
static inline float om_acos(float x) {
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return u2f(0x7fc00000u);
    const float s = std::sqrt((1.0f - x) * (1.0f + x));
    return om_atan2(s, x);
}

} // namespace ora
