// oracle/twin_math.hpp — scalar math of the CPU twin.
//
// TEST INFRASTRUCTURE: part of the oracle (the CPU restatement of rayna_engine's render path).  Only
// tests/, __graft_entry__.smoke() and bench.py's CPU baseline may load it; the product never does.
//
// Double precision uses the C library, as the engine's `Number = f64` math does
// (rayna_engine/src/core/types.rs:7).  Single precision uses functions written here from published
// methods, with a fixed operation order and no fused multiply-add, so that a GPU kernel that evaluates
// the same formulas in the same order produces the same bits:
//   sin/cos of 2*pi*r   exact quadrant reduction in turns + Taylor series of degree 9/10 on [-pi/4, pi/4]
//   log                 exponent split + atanh series  log(m) = 2*atanh((m-1)/(m+1))   (Abramowitz & Stegun 4.1.27)
//   exp                 Cody-Waite reduction by ln 2 (two-part constant) + Taylor series of degree 7
//   cbrt                bit-level initial estimate + four Newton steps y <- (2y + x/y^2)/3
// Every function is checked against the C library's double-precision result in tests/test_oracle_math.py.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace twin {

inline float f32_from_bits(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t f32_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// ---- single precision -------------------------------------------------------------------------

// sin and cos of 2*pi*r for r in [0, 1) (the angles the engine's samplers draw, shared/rng.rs:75-87).
// q = 4r is exact; k = nearest integer; f = q - k in [-1/2, 1/2] is exact (Sterbenz); angle = pi/2*(k+f).
inline void sincos_turns_f32(float r, float *s_out, float *c_out) {
    const float q = r * 4.0f;
    const float k = std::rint(q);
    const float f = q - k;
    const float f2 = f * f;
    // sin(pi/2 f) = f*(S1 + f2*(S3 + f2*(S5 + f2*(S7 + f2*S9))))
    const float S1 = 0x1.921fb6p+0f, S3 = -0x1.4abbcep-1f, S5 = 0x1.466bc6p-4f, S7 = -0x1.32d2ccp-8f,
                S9 = 0x1.507834p-13f;
    // cos(pi/2 f) = 1 + f2*(C2 + f2*(C4 + f2*(C6 + f2*(C8 + f2*C10))))
    const float C2 = -0x1.3bd3ccp+0f, C4 = 0x1.03c1f0p-2f, C6 = -0x1.55d3c8p-6f, C8 = 0x1.e1f506p-11f,
                C10 = -0x1.a6d1f2p-16f;
    const float sp = f * (S1 + f2 * (S3 + f2 * (S5 + f2 * (S7 + f2 * S9))));
    const float cp = 1.0f + f2 * (C2 + f2 * (C4 + f2 * (C6 + f2 * (C8 + f2 * C10))));
    const int quadrant = ((int)k) & 3;
    float s, c;
    switch (quadrant) {
        case 0: s = sp; c = cp; break;
        case 1: s = cp; c = -sp; break;
        case 2: s = -sp; c = -cp; break;
        default: s = -cp; c = sp; break;
    }
    *s_out = s;
    *c_out = c;
}

// Natural logarithm.  x = m * 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
// |s| <= 0.1716, series to s^11 in Horner form with fmas (as the device evaluates it).  ln 2 is split in a 16-bit
// head and a tail so e*ln2_hi is exact.
inline float log_f32(float x) {
    if (x != x || x < 0.0f) return std::numeric_limits<float>::quiet_NaN();
    if (x == 0.0f) return -std::numeric_limits<float>::infinity();
    if (x == std::numeric_limits<float>::infinity()) return x;
    int e_adj = 0;
    if (x < 0x1.0p-126f) {  // subnormal: scale into the normal range first
        x = x * 0x1.0p+24f;
        e_adj = -24;
    }
    const uint32_t b = f32_bits(x);
    int e = (int)((b >> 23) & 0xffu) - 127 + e_adj;
    float m = f32_from_bits((b & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 0x1.6a09e6p+0f) {                                  // > sqrt(2)
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float L3 = 0x1.555556p-2f, L5 = 0x1.99999ap-3f, L7 = 0x1.24924ap-3f, L9 = 0x1.c71c72p-4f,
                L11 = 0x1.745d18p-4f;
    const float poly = s2 * std::fma(s2, std::fma(s2, std::fma(s2, std::fma(s2, L11, L9), L7), L5), L3);
    const float log_m = std::fma(s, poly, s) * 2.0f;
    const float fe = (float)e;
    const float LN2_HI = 0x1.62e4p-1f, LN2_LO = 0x1.7f7d1cp-20f;
    return std::fma(fe, LN2_HI, std::fma(fe, LN2_LO, log_m));
}

// Exponential.  k = rint(x / ln 2), r = x - k ln2 (Cody-Waite, two parts), exp(r) by Taylor to r^7,
// then scaled by 2^k in two exact steps.
inline float exp_f32(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return std::numeric_limits<float>::infinity();
    if (x < -103.97208404541016f) return 0.0f;
    const float LOG2E = 0x1.715476p+0f;
    const float LN2_HI = 0x1.62e4p-1f, LN2_LO = 0x1.7f7d1cp-20f;
    const float k = std::rint(x * LOG2E);
    const float r = (x - k * LN2_HI) - k * LN2_LO;
    const float E2 = 0x1.0p-1f, E3 = 0x1.555556p-3f, E4 = 0x1.555556p-5f, E5 = 0x1.111112p-7f,
                E6 = 0x1.6c16c2p-10f, E7 = 0x1.a01a02p-13f;
    const float p = 1.0f + r * (1.0f + r * (E2 + r * (E3 + r * (E4 + r * (E5 + r * (E6 + r * E7))))));
    int ki = (int)k;
    // 2^k as the product of two powers of two that are both normal floats
    const int k1 = ki / 2;
    const int k2 = ki - k1;
    const float s1 = f32_from_bits((uint32_t)(k1 + 127) << 23);
    const float s2 = f32_from_bits((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// x^y for the UI's gamma (x >= 0, y > 0): exp(y log x); 0 for x = 0, NaN for x < 0 or NaN.
inline float pow_f32(float x, float y) {
    if (x != x || x < 0.0f) return std::numeric_limits<float>::quiet_NaN();
    if (x == 0.0f) return 0.0f;
    return exp_f32(y * log_f32(x));
}

// Cube root of x >= 0: the classic exponent-thirds initial guess (within ~6% over the normal range) and four
// Newton steps (the radius of a uniform point in the unit ball; the distance render mode).
inline float cbrt01_f32(float x) {
    if (x <= 0.0f) return 0.0f;
    float y = f32_from_bits(f32_bits(x) / 3u + 0x2a5137a0u);
    for (int i = 0; i < 4; ++i) y = (y + y + x / (y * y)) * (1.0f / 3.0f);
    return y;
}

// sin and cos of x radians, |x| < 2^10 (the Mandelbulb angles, at most power * pi).
// Cody-Waite reduction by pi/2 in three parts (head and middle carry 11 significant bits, so k*part is
// exact for |k| < 2^12), then Taylor series of degree 9 (sin) and 10 (cos) on [-pi/4, pi/4].
inline void sincos_rad_f32(float x, float *s_out, float *c_out) {
    const float TWO_OVER_PI = 0x1.45f306p-1f;
    const float P1 = 0x1.92p+0f, P2 = 0x1.fb4p-12f, P3 = 0x1.4442d2p-24f;
    const float k = std::rint(x * TWO_OVER_PI);
    const float r = ((x - k * P1) - k * P2) - k * P3;
    const float r2 = r * r;
    const float S3 = -0x1.555556p-3f, S5 = 0x1.111112p-7f, S7 = -0x1.a01a02p-13f, S9 = 0x1.71de3ap-19f;
    const float C2 = -0x1.0p-1f, C4 = 0x1.555556p-5f, C6 = -0x1.6c16c2p-10f, C8 = 0x1.a01a02p-16f,
                C10 = -0x1.27e4fcp-22f;
    const float sp = r + r * (r2 * (S3 + r2 * (S5 + r2 * (S7 + r2 * S9))));
    const float cp = 1.0f + r2 * (C2 + r2 * (C4 + r2 * (C6 + r2 * (C8 + r2 * C10))));
    const int quadrant = ((int)k) & 3;
    float s, c;
    switch (quadrant) {
        case 0: s = sp; c = cp; break;
        case 1: s = cp; c = -sp; break;
        case 2: s = -sp; c = -cp; break;
        default: s = -cp; c = sp; break;
    }
    *s_out = s;
    *c_out = c;
}

// atan of z, |z| <= tan(pi/8): Taylor series to z^21.
inline float atan_small_f32(float z) {
    const float z2 = z * z;
    const float A3 = -0x1.555556p-2f, A5 = 0x1.99999ap-3f, A7 = -0x1.24924ap-3f, A9 = 0x1.c71c72p-4f,
                A11 = -0x1.745d18p-4f, A13 = 0x1.3b13b2p-4f, A15 = -0x1.111112p-4f, A17 = 0x1.e1e1e2p-5f,
                A19 = -0x1.af286cp-5f, A21 = 0x1.861862p-5f;
    const float p =
        z2 * (A3 + z2 * (A5 + z2 * (A7 + z2 * (A9 + z2 * (A11 + z2 * (A13 + z2 * (A15 + z2 * (A17 + z2 * (A19 + z2 * A21)))))))));
    return z + z * p;
}

// atan2(y, x): octant reduction (Cephes-style: t > tan(3pi/8) -> pi/2 - atan(1/t); t > tan(pi/8) ->
// pi/4 + atan((t-1)/(t+1))), then the quadrant from the signs of x and y.
inline float atan2_f32(float y, float x) {
    if (x != x || y != y) return x + y;
    const float PI = 0x1.921fb6p+1f, PI_2 = 0x1.921fb6p+0f, PI_4 = 0x1.921fb6p-1f;
    const float ax = std::fabs(x), ay = std::fabs(y);
    if (ax == 0.0f && ay == 0.0f) {
        const float base = std::signbit(x) ? PI : 0.0f;
        return std::signbit(y) ? -base : base;
    }
    float a;
    if (ay <= ax) {
        const float t = ay / ax;
        if (t > 0x1.a8279ap-2f) a = PI_4 + atan_small_f32((t - 1.0f) / (t + 1.0f));
        else a = atan_small_f32(t);
    } else {
        const float t = ax / ay;  // atan(ay/ax) = pi/2 - atan(ax/ay)
        if (t > 0x1.a8279ap-2f) a = PI_4 - atan_small_f32((t - 1.0f) / (t + 1.0f));
        else a = PI_2 - atan_small_f32(t);
    }
    if (std::signbit(x)) a = PI - a;
    return std::signbit(y) ? -a : a;
}

// acos(x) = atan2(sqrt((1 - x)(1 + x)), x) for x in [-1, 1].
inline float acos_f32(float x) {
    const float s = std::sqrt((1.0f - x) * (1.0f + x));
    return atan2_f32(s, x);
}

// ---- double precision: the C library -----------------------------------------------------------

inline void sincos_turns_f64(double r, double *s, double *c) {
    // rng::normal_on_unit_sphere evaluates cos(2*PI*r1) and sin(2*PI*r1) (shared/rng.rs:91-92)
    const double a = 2.0 * M_PI * r;
    *s = std::sin(a);
    *c = std::cos(a);
}

// Precision-dispatched helpers used by the templated twin.
template <class R>
struct Math;

template <>
struct Math<float> {
    static float sqrt(float x) { return std::sqrt(x); }
    static void sincos_turns(float r, float *s, float *c) { sincos_turns_f32(r, s, c); }
    static float log(float x) { return log_f32(x); }
    static float exp(float x) { return exp_f32(x); }
    static float cbrt01(float x) { return cbrt01_f32(x); }
    static float fmin(float a, float b) { return std::fmin(a, b); }
    static float fmax(float a, float b) { return std::fmax(a, b); }
    static float big() { return std::numeric_limits<float>::max(); }
    static void sincos(float x, float *s, float *c) { sincos_rad_f32(x, s, c); }
    static float atan2(float y, float x) { return atan2_f32(y, x); }
    static float acos(float x) { return acos_f32(x); }
};

template <>
struct Math<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static void sincos_turns(double r, double *s, double *c) { sincos_turns_f64(r, s, c); }
    static double log(double x) { return std::log(x); }
    static double exp(double x) { return std::exp(x); }
    static double cbrt01(double x) { return std::cbrt(x); }
    static double fmin(double a, double b) { return std::fmin(a, b); }
    static double fmax(double a, double b) { return std::fmax(a, b); }
    static double big() { return std::numeric_limits<double>::max(); }
    static void sincos(double x, double *s, double *c) {
        *s = std::sin(x);
        *c = std::cos(x);
    }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double acos(double x) { return std::acos(x); }
};

// Rust's f64::signum: 1 for +0 and positive values, -1 for -0 and negative values, NaN for NaN.
template <class R>
inline R signum(R x) {
    if (x != x) return x;
    return std::signbit(x) ? R(-1) : R(1);
}

}  // namespace twin
