// Single-precision math and counter-based RNG of the gfx950 kernels.
//
// Every function is written out with a fixed operation order and compiled without floating-point contraction
// (-ffp-contract=off) and with correctly rounded division and square root, so the GPU evaluates exactly the
// IEEE operations the formulas name:
//   sin/cos(2 pi r)   exact quadrant reduction in turns, Taylor series of degree 9 / 10 on [-pi/4, pi/4]
//   sin/cos(x)        Cody-Waite reduction by pi/2 in three parts, the same series
//   log               exponent split, 2 atanh((m-1)/(m+1)) series to s^11 (Abramowitz & Stegun 4.1.27)
//   exp               Cody-Waite reduction by ln 2, Taylor series to r^7, exact scaling by 2^k
//   cbrt              bit-level estimate and four Newton steps y <- (2y + x/y^2)/3
//   atan2 / acos      octant reduction, Taylor series of atan to z^21; acos(x) = atan2(sqrt((1-x)(1+x)), x)
//   Philox4x32-10     Salmon et al., SC 2011 (counter-based RNG)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rayna {
namespace dm {

__device__ __forceinline__ void sincos_turns(float r, float *s_out, float *c_out) {
    const float q = r * 4.0f;
    const float k = rintf(q);
    const float f = q - k;
    const float f2 = f * f;
    const float S1 = 0x1.921fb6p+0f, S3 = -0x1.4abbcep-1f, S5 = 0x1.466bc6p-4f, S7 = -0x1.32d2ccp-8f,
                S9 = 0x1.507834p-13f;
    const float C2 = -0x1.3bd3ccp+0f, C4 = 0x1.03c1f0p-2f, C6 = -0x1.55d3c8p-6f, C8 = 0x1.e1f506p-11f,
                C10 = -0x1.a6d1f2p-16f;
    const float sp = f * (S1 + f2 * (S3 + f2 * (S5 + f2 * (S7 + f2 * S9))));
    const float cp = 1.0f + f2 * (C2 + f2 * (C4 + f2 * (C6 + f2 * (C8 + f2 * C10))));
    // quadrant q: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp), by selects and exact negations
    const int quadrant = ((int)k) & 3;
    const float s = (quadrant & 1) ? cp : sp, c = (quadrant & 1) ? sp : cp;
    *s_out = (quadrant & 2) ? -s : s;
    *c_out = ((quadrant + 1) & 2) ? -c : c;
}

__device__ __forceinline__ float log_f(float x) {
    if (x != x || x < 0.0f) return __uint_as_float(0x7fc00000u);
    if (x == 0.0f) return -__builtin_inff();
    if (x == __builtin_inff()) return x;
    int e_adj = 0;
    if (x < 0x1.0p-126f) {
        x = x * 0x1.0p+24f;
        e_adj = -24;
    }
    const uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xffu) - 127 + e_adj;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float L3 = 0x1.555556p-2f, L5 = 0x1.99999ap-3f, L7 = 0x1.24924ap-3f, L9 = 0x1.c71c72p-4f,
                L11 = 0x1.745d18p-4f;
    const float poly = s2 * fmaf(s2, fmaf(s2, fmaf(s2, fmaf(s2, L11, L9), L7), L5), L3);
    const float log_m = fmaf(s, poly, s) * 2.0f;
    const float fe = (float)e;
    const float LN2_HI = 0x1.62e4p-1f, LN2_LO = 0x1.7f7d1cp-20f;
    return fmaf(fe, LN2_HI, fmaf(fe, LN2_LO, log_m));
}

__device__ __forceinline__ float exp_f(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97208404541016f) return 0.0f;
    const float LOG2E = 0x1.715476p+0f;
    const float LN2_HI = 0x1.62e4p-1f, LN2_LO = 0x1.7f7d1cp-20f;
    const float k = rintf(x * LOG2E);
    const float r = (x - k * LN2_HI) - k * LN2_LO;
    const float E2 = 0x1.0p-1f, E3 = 0x1.555556p-3f, E4 = 0x1.555556p-5f, E5 = 0x1.111112p-7f,
                E6 = 0x1.6c16c2p-10f, E7 = 0x1.a01a02p-13f;
    const float p = 1.0f + r * (1.0f + r * (E2 + r * (E3 + r * (E4 + r * (E5 + r * (E6 + r * E7))))));
    const int ki = (int)k;
    const int k1 = ki / 2;
    const int k2 = ki - k1;
    const float s1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    const float s2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// x^y for the UI's gamma (x >= 0, y > 0): exp(y log x); 0 for x = 0, NaN for x < 0 or NaN.
__device__ __forceinline__ float pow_f(float x, float y) {
    if (x != x || x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return 0.0f;
    return exp_f(y * log_f(x));
}

// Cube root of x >= 0: exponent-thirds initial guess and four Newton steps (also valid above 1).
__device__ __forceinline__ float cbrt01(float x) {
    if (x <= 0.0f) return 0.0f;
    float y = __uint_as_float(__float_as_uint(x) / 3u + 0x2a5137a0u);
    for (int i = 0; i < 4; ++i) y = (y + y + x / (y * y)) * (1.0f / 3.0f);
    return y;
}

__device__ __forceinline__ void sincos_rad(float x, float *s_out, float *c_out) {
    const float TWO_OVER_PI = 0x1.45f306p-1f;
    const float P1 = 0x1.92p+0f, P2 = 0x1.fb4p-12f, P3 = 0x1.4442d2p-24f;
    const float k = rintf(x * TWO_OVER_PI);
    const float r = ((x - k * P1) - k * P2) - k * P3;
    const float r2 = r * r;
    const float S3 = -0x1.555556p-3f, S5 = 0x1.111112p-7f, S7 = -0x1.a01a02p-13f, S9 = 0x1.71de3ap-19f;
    const float C2 = -0x1.0p-1f, C4 = 0x1.555556p-5f, C6 = -0x1.6c16c2p-10f, C8 = 0x1.a01a02p-16f,
                C10 = -0x1.27e4fcp-22f;
    const float sp = r + r * (r2 * (S3 + r2 * (S5 + r2 * (S7 + r2 * S9))));
    const float cp = 1.0f + r2 * (C2 + r2 * (C4 + r2 * (C6 + r2 * (C8 + r2 * C10))));
    // quadrant q: (s, c) = (sp, cp), (cp, -sp), (-sp, -cp), (-cp, sp), by selects and exact negations
    const int quadrant = ((int)k) & 3;
    const float s = (quadrant & 1) ? cp : sp, c = (quadrant & 1) ? sp : cp;
    *s_out = (quadrant & 2) ? -s : s;
    *c_out = ((quadrant + 1) & 2) ? -c : c;
}

__device__ __forceinline__ float atan_small(float z) {
    const float z2 = z * z;
    const float A3 = -0x1.555556p-2f, A5 = 0x1.99999ap-3f, A7 = -0x1.24924ap-3f, A9 = 0x1.c71c72p-4f,
                A11 = -0x1.745d18p-4f, A13 = 0x1.3b13b2p-4f, A15 = -0x1.111112p-4f, A17 = 0x1.e1e1e2p-5f,
                A19 = -0x1.af286cp-5f, A21 = 0x1.861862p-5f;
    const float p =
        z2 * (A3 + z2 * (A5 + z2 * (A7 + z2 * (A9 + z2 * (A11 + z2 * (A13 + z2 * (A15 + z2 * (A17 + z2 * (A19 + z2 * A21)))))))));
    return z + z * p;
}

__device__ __forceinline__ float atan2_f(float y, float x) {
    if (x != x || y != y) return x + y;
    const float PI = 0x1.921fb6p+1f, PI_2 = 0x1.921fb6p+0f, PI_4 = 0x1.921fb6p-1f;
    const float ax = fabsf(x), ay = fabsf(y);
    if (ax == 0.0f && ay == 0.0f) {
        const float base = signbit(x) ? PI : 0.0f;
        return signbit(y) ? -base : base;
    }
    // Four reductions: t = ay / ax or ax / ay (the smaller over the larger), shifted by (t - 1) / (t + 1) when t is
    // large, then a = p, PI_4 + p, PI_4 - p or PI_2 - p.  They share one quotient, one shifted quotient and one
    // polynomial, chosen by selects, so a wave whose lanes take different reductions pays for one.
    const bool swap = !(ay <= ax);
    const float t = (swap ? ax : ay) / (swap ? ay : ax);
    const bool big = t > 0x1.a8279ap-2f;
    const float p = atan_small(big ? (t - 1.0f) / (t + 1.0f) : t);
    float a = swap ? (big ? PI_4 - p : PI_2 - p) : (big ? PI_4 + p : p);
    if (signbit(x)) a = PI - a;
    return signbit(y) ? -a : a;
}

__device__ __forceinline__ float acos_f(float x) {
    const float s = sqrtf((1.0f - x) * (1.0f + x));
    return atan2_f(s, x);
}

// Rust's f32::signum: +-1 by the sign bit, NaN for NaN.
__device__ __forceinline__ float signum(float x) {
    if (x != x) return x;
    return signbit(x) ? -1.0f : 1.0f;
}

// ---- Philox4x32-10 ----

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * 0x1.0p-24f; }

// Draw-site slots (the fourth counter word); identical to the CPU restatement's.
enum : uint32_t { SLOT_CAMERA = 0, SLOT_SCATTER = 1, SLOT_MEDIUM_AUX = 2, SLOT_MEDIUM_UV = 3, SLOT_MEDIUM_BASE = 16 };

struct Key {
    uint32_t k0, k1;
};

__device__ __forceinline__ void draw4(const Key &key, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t slot,
                                      float u[4]) {
    const U4 o = philox(pixel, sample, bounce, slot, key.k0, key.k1);
    u[0] = u01(o.x);
    u[1] = u01(o.y);
    u[2] = u01(o.z);
    u[3] = u01(o.w);
}

}  // namespace dm
}  // namespace rayna
