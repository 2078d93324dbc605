// oracle/twin_rng.hpp — counter-based random numbers and samplers of the CPU twin.
//
// TEST INFRASTRUCTURE: part of the oracle; only tests/, smoke() and bench.py's CPU baseline load it.
//
// The engine draws from entropy-seeded generators pooled per worker (rayna_engine/src/render/renderer.rs:
// 191-216) and uses rejection loops for points in the unit ball and disk (shared/rng.rs:64-72,134-142), so
// no run of it can be repeated.  The twin (and the GPU path) replace both:
//  * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011),
//    key = (seed low word, seed high word ^ frame), counter = (pixel, sample, bounce, slot): each draw site
//    has its own stream, independent of the order in which paths or primitives are processed;
//  * closed-form samplers with the same distributions as the rejection loops:
//    disk = sqrt(u) * (cos 2pi v, sin 2pi v); ball = cbrt(u) * normal_on_unit_sphere(v, w).
#pragma once

#include <cstdint>

#include "twin_math.hpp"

namespace twin {

// Draw-site slots (the fourth counter word).
enum : uint32_t {
    SLOT_CAMERA = 0,        // lanes 0,1: pixel jitter (renderer.rs:377,388); lanes 2,3: lens disk (camera.rs:208)
    SLOT_SCATTER = 1,       // the material's scatter draws at this bounce (material/*.rs)
    SLOT_MEDIUM_AUX = 2,    // the volume hit's arbitrary normals (object/volumetric.rs:151-153)
    SLOT_MEDIUM_UV = 3,     // lanes 0,1: the volume hit's arbitrary uv (object/volumetric.rs:153)
    SLOT_MEDIUM_BASE = 16,  // + object index: free-flight distance of that volume (object/volumetric.rs:128)
};

struct Philox4x32 {
    static constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    static constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;

    static void round(uint32_t c[4], const uint32_t k[2]) {
        const uint64_t p0 = (uint64_t)M0 * c[0];
        const uint64_t p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k[0];
        const uint32_t n2 = hi0 ^ c[3] ^ k[1];
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }

    static void generate(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
        uint32_t k[2] = {key[0], key[1]};
        round(c, k);
        for (int r = 1; r < 10; ++r) {
            k[0] += W0;
            k[1] += W1;
            round(c, k);
        }
        for (int i = 0; i < 4; ++i) out[i] = c[i];
    }
};

// Uniform in [0, 1) with 24 random bits: exactly representable in float and double alike, so the fp32
// and fp64 twins (and the GPU) see the same numbers.
inline float u01_f32(uint32_t x) { return (float)(x >> 8) * 0x1.0p-24f; }

struct RngKey {
    uint32_t k[2];
};

inline RngKey make_key(uint64_t seed, uint32_t frame) {
    RngKey key;
    key.k[0] = (uint32_t)seed;
    key.k[1] = (uint32_t)(seed >> 32) ^ frame;
    return key;
}

// Four uniforms of one draw site.
template <class R>
inline void draw4(const RngKey &key, uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t slot, R u[4]) {
    const uint32_t ctr[4] = {pixel, sample, bounce, slot};
    uint32_t o[4];
    Philox4x32::generate(ctr, key.k, o);
    for (int i = 0; i < 4; ++i) u[i] = (R)u01_f32(o[i]);
}

template <class R>
struct V3;  // defined in twin.hpp

}  // namespace twin
