// Wavefront path-tracing kernels for gfx950 (CDNA4, wave64): generate -> (extend -> shade)* -> resolve ->
// finalize.  Each kernel restates one stage of rayna_engine's per-pixel recursion
// (rayna_engine/src/render/renderer.rs:350-572) for a whole chunk of paths held as SoA queues in HBM.
#include <hip/hip_runtime.h>
#include <float.h>

#include <algorithm>
#include <mutex>
#include <stdint.h>

#include "../../include/rayna_gpu.h"
#include "device_math.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace rayna {

namespace {

constexpr int BLOCK = 256;
// Extend-kernel shape (overridable at build time for tuning sweeps).
#ifndef RAYNA_EXT_BLOCK
#define RAYNA_EXT_BLOCK 128
#endif
#ifndef RAYNA_EXT_LDS_STACK
#define RAYNA_EXT_LDS_STACK 11
#endif
constexpr int EXT_BLOCK = RAYNA_EXT_BLOCK;
constexpr int EXT_LDS_STACK = RAYNA_EXT_LDS_STACK;  // traversal stack entries per lane held in LDS
constexpr float T_MIN = 1e-3f;                // render interval start (renderer.rs:239)
constexpr float BOX_SLACK = 1.0f + 0x1.0p-16f;  // conservative slab test: culling never drops the closest hit

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 neg(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }

// sqrtf(x), correctly rounded, without the pass-through of zeros and infinities that sqrtf's general sequence ends
// with: inputs below 2^-96 are scaled by 2^32 (the root by 2^-16 back), the hardware square root is corrected by
// the signs of the residuals x - y'y at the root's neighbours y' = y -+ 1 ulp, and for zeros and infinities that
// correction already leaves the root as it is.  Equal to sqrtf bit for bit on every input (checked over all 2^32
// bit patterns: rayna_debug_sqrt_mismatches).  For the SDF sphere distances, which the march evaluates sixteen at
// a time.
__device__ __forceinline__ float sqrt_rn_corrected(float xs) {  // the hardware root of xs, corrected
    const float y = __builtin_amdgcn_sqrtf(xs);
    const float ym = __uint_as_float(__float_as_uint(y) - 1u), yp = __uint_as_float(__float_as_uint(y) + 1u);
    float r = fmaf(-ym, y, xs) <= 0.0f ? ym : y;
    return fmaf(-yp, y, xs) > 0.0f ? yp : r;
}
__device__ __forceinline__ float sqrt_rn(float x) {
    const bool tiny = x < 0x1p-96f;
    const float r = sqrt_rn_corrected(tiny ? x * 0x1p32f : x);
    return tiny ? r * 0x1p-16f : r;
}
// sqrt_rn with the scaling taken only by waves that have an input below 2^-96 (a point within 2^-48 of a sphere's
// centre): the same value on every input.
__device__ __forceinline__ float sqrt_rn_wave(float x) {
    if (__any(x < 0x1p-96f)) return sqrt_rn(x);
    return sqrt_rn_corrected(x);
}
// |a| from the fused sum of squares fma(z, z, fma(y, y, x x)) and a correctly rounded square root: the SDF sphere
// distances, as the oracle's single precision takes them.
__device__ __forceinline__ float length_fma(f3 a) {
    return sqrt_rn_wave(fmaf(a.z, a.z, fmaf(a.y, a.y, a.x * a.x)));
}
__device__ __forceinline__ f3 normalize(f3 a) { return a * (1.0f / length(a)); }
__device__ __forceinline__ bool try_normalize(f3 a, f3 *out) {  // glam try_normalize
    const float rcp = 1.0f / length(a);
    if (isfinite(rcp) && rcp > 0.0f) {
        *out = a * rcp;
        return true;
    }
    return false;
}
__device__ __forceinline__ f3 reflect(f3 d, f3 n) { return d - n * (2.0f * dot(d, n)); }  // shared/math.rs:26
__device__ __forceinline__ f3 refract(f3 v, f3 n, float ratio) {                        // shared/math.rs:20-25
    const float cos_theta = fminf(dot(neg(v), n), 1.0f);
    const f3 perp = (v + n * cos_theta) * ratio;
    const f3 par = n * -sqrtf(fabsf(1.0f - dot(perp, perp)));
    return perp + par;
}

struct Col {
    float r, g, b;
};
__device__ __forceinline__ Col operator+(Col a, Col b) { return {a.r + b.r, a.g + b.g, a.b + b.b}; }
__device__ __forceinline__ Col operator*(Col a, Col b) { return {a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ Col operator*(Col a, float s) { return {a.r * s, a.g * s, a.b * s}; }

struct DRay {  // shared/ray.rs:15-23
    f3 o, d, inv;
};
__device__ __forceinline__ DRay make_ray(f3 o, f3 d) {
    DRay r;
    r.o = o;
    r.d = normalize(d);
    r.inv = f3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    return r;
}
__device__ __forceinline__ DRay ray_from_dir(f3 o, f3 unit_d) {
    DRay r;
    r.o = o;
    r.d = unit_d;
    r.inv = f3{1.0f / unit_d.x, 1.0f / unit_d.y, 1.0f / unit_d.z};
    return r;
}
// A ray whose reciprocal direction is left unset (shading needs it only for box hits: with_inv).
__device__ __forceinline__ DRay ray_lazy(f3 o, f3 unit_d) { return DRay{o, unit_d, f3{0.0f, 0.0f, 0.0f}}; }
__device__ __forceinline__ DRay with_inv(const DRay &r) {
    return DRay{r.o, r.d, f3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z}};
}
__device__ __forceinline__ f3 at(const DRay &r, float t) { return r.o + r.d * t; }

// ------------------------------------------------------------------------------------------ primitives

// SphereMesh::intersect (sphere.rs:44-99): the engine's quadratic, evaluated in double from the single-precision
// centre, radius and ray, keeping a = |d|^2.  In single precision c = |oc|^2 - r^2 and b^2 - c cancel (the
// r = 1000 ground sphere's roots would be off by up to 5e-4 relative), and the engine's a = 1 (sphere.rs:50-52)
// holds only to ~6e-8 for a single-precision unit direction, which c / (2 sqrt(disc)) magnifies to ~1e-4 for a
// small, distant sphere.  In double o - c is exact and the rest rounds far below single precision; the roots are
// (-b -+ sqrt(b^2 - a c)) / a, the division taken as the product with 2 - a (exact to (1 - a)^2 ~ 1e-14).  The
// root, rounded once to single precision, stays within 2e-6 of the double-precision engine's; each root is
// checked against [t0, t1] after that rounding.
__device__ __forceinline__ bool sphere_t(const GSphere &s, const DRay &r, float t0, float t1, float *t) {
    const double ox = (double)r.o.x - (double)s.cx, oy = (double)r.o.y - (double)s.cy,
                 oz = (double)r.o.z - (double)s.cz;
    const double dx = r.d.x, dy = r.d.y, dz = r.d.z;
    const double r2 = (double)s.radius * (double)s.radius;
    const double half_b = (ox * dx + oy * dy) + oz * dz;
    const double cc = ((ox * ox + oy * oy) + oz * oz) - r2;
    const double a = (dx * dx + dy * dy) + dz * dz;
    const double disc = (half_b * half_b) - a * cc;
    if (disc < 0.0) return false;
    const double sq = sqrt(disc);
    const double inv_a = 2.0 - a;
    float root = (float)((-half_b - sq) * inv_a);
    if (!(t0 <= root && root <= t1)) {
        root = (float)((-half_b + sq) * inv_a);
        if (!(t0 <= root && root <= t1)) return false;
    }
    *t = root;
    return true;
}

// Triangle::intersect (triangle.rs:53-109), Moller-Trumbore; e1, e2 precomputed in float.
__device__ __forceinline__ bool tri_t(const GTri &g, const DRay &r, float t0, float t1, float *t, float *uu, float *vv,
                                      float *det_out) {
    const f3 v0{g.v0x, g.v0y, g.v0z}, e1{g.e1x, g.e1y, g.e1z}, e2{g.e2x, g.e2y, g.e2z};
    const f3 p = cross(r.d, e2);
    const float det = dot(e1, p);
    if (det == 0.0f) return false;
    const float inv_det = 1.0f / det;
    const f3 tv = r.o - v0;
    const float u = dot(tv, p) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(tv, e1);
    const float v = dot(r.d, q) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    const float tt = dot(e2, q) * inv_det;
    if (!(t0 <= tt && tt <= t1)) return false;
    *t = tt;
    *uu = u;
    *vv = v;
    *det_out = det;
    return true;
}

// Triangle::interpolate_normals (triangle.rs:114-119): fold from zero, then try_normalize.
__device__ __forceinline__ bool tri_normal(const GTriNormals &n, float u, float v, f3 *out) {
    const float b0 = (1.0f - u) - v;
    const f3 s{((0.0f + n.n0x * b0) + n.n1x * u) + n.n2x * v, ((0.0f + n.n0y * b0) + n.n1y * u) + n.n2y * v,
               ((0.0f + n.n0z * b0) + n.n1z * u) + n.n2z * v};
    return try_normalize(s, out);
}

// Planar::intersect_bounded + ParallelogramMesh::intersect (planar/mod.rs:161-194, parallelogram.rs:49-57)
__device__ __forceinline__ bool para_t(const GPara &g, const DRay &r, float t0, float t1, float *t,
                                       float *uv_out = nullptr) {
    const f3 n{g.nx, g.ny, g.nz};
    const float denom = dot(n, r.d);
    if (denom == 0.0f) return false;
    const float tt = -(dot(n, r.o) + g.d) / denom;
    if (!(t0 <= tt && tt <= t1)) return false;
    const f3 pos = at(r, tt);
    const f3 pl = pos - f3{g.px, g.py, g.pz};
    const f3 w{g.wx, g.wy, g.wz};
    const float alpha = dot(w, cross(pl, f3{g.vx, g.vy, g.vz}));
    const float beta = dot(w, cross(f3{g.ux, g.uy, g.uz}, pl));
    if (!(alpha <= 1.0f && beta <= 1.0f && alpha >= 0.0f && beta >= 0.0f)) return false;
    *t = tt;
    if (uv_out) {  // planar/mod.rs:191
        uv_out[0] = alpha;
        uv_out[1] = beta;
    }
    return true;
}

// AxisBoxMesh::intersect (axis_box.rs:64-145).  axis_out: the face axis hit; winding/sgn for the normals.
__device__ __forceinline__ bool box_t(const GBox &g, const DRay &r, float t0, float t1, float *t, int *axis_out,
                                      float *winding_out, f3 *sgn_out, float *uv_out = nullptr) {
    const f3 ro = r.o - f3{g.cx, g.cy, g.cz};
    const f3 rd = r.d;
    const f3 radius{g.rx, g.ry, g.rz};
    const float ax = fabsf(ro.x) * g.ix, ay = fabsf(ro.y) * g.iy, az = fabsf(ro.z) * g.iz;
    const float winding = dm::signum(fmaxf(fmaxf(ax, ay), az) - 1.0f);
    const f3 sgn{-dm::signum(rd.x), -dm::signum(rd.y), -dm::signum(rd.z)};
    f3 pd = mul(radius * winding, sgn) - ro;
    pd = mul(pd, r.inv);
    const float pdv[3] = {pd.x, pd.y, pd.z};
    const float rov[3] = {ro.x, ro.y, ro.z};
    const float rdv[3] = {rd.x, rd.y, rd.z};
    const float rad[3] = {radius.x, radius.y, radius.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const float dist = pdv[a];
        if (!(t0 <= dist && dist <= t1)) continue;
        const float ub = rov[b] + rdv[b] * dist, uc = rov[c] + rdv[c] * dist;
        if (fabsf(ub) < rad[b] && fabsf(uc) < rad[c]) {
            *t = dist;
            *axis_out = a;
            *winding_out = winding;
            *sgn_out = sgn;
            if (uv_out) {  // remapped from [-radius, radius] to [0, 1] (axis_box.rs:121-122)
                uv_out[0] = (ub / rad[b] + 1.0f) / 2.0f;
                uv_out[1] = (uc / rad[c] + 1.0f) / 2.0f;
            }
            return true;
        }
    }
    return false;
}

// SPH: the scene intersects spheres in traversal (without, the double-precision sphere test is compiled out: a
// sphere boundary then never occurs).
template <bool SPH>
__device__ __forceinline__ bool boundary_t(const SceneView &sv, const GVolume &v, const DRay &r, float t0, float t1,
                                           float *t) {
    if (v.boundary_kind == ITEM_SPHERE) return SPH && sphere_t(sv.spheres[v.boundary_index], r, t0, t1, t);
    int axis;
    float w;
    f3 sg;
    return box_t(sv.boxes[v.boundary_index], r, t0, t1, t, &axis, &w, &sg);
}

// VolumetricObject::full_intersect (object/volumetric.rs:86-160): entry with an unbounded interval, exit after
// entry + 0.001, free flight -ln(U)/density.
template <bool SPH>
__device__ __forceinline__ bool volume_t(const SceneView &sv, const GVolume &v, const DRay &r, float t_end,
                                         const dm::Key &key, uint32_t pixel, uint32_t sample, uint32_t bounce,
                                         float *t) {
    float entry;
    if (!boundary_t<SPH>(sv, v, r, -__builtin_inff(), __builtin_inff(), &entry)) return false;
    entry = fmaxf(entry, T_MIN);
    float exit_d;
    if (!boundary_t<SPH>(sv, v, r, entry + 0.001f, __builtin_inff(), &exit_d)) return false;
    const float inside = exit_d - entry;
    float u[4];
    dm::draw4(key, pixel, sample, bounce, dm::SLOT_MEDIUM_BASE + (uint32_t)v.object, u);
    const float hit_dist = v.neg_inv_density * dm::log_f(u[0]);
    const float dist = entry + hit_dist;
    if (hit_dist > inside) return false;
    if (dist > t_end) return false;
    *t = dist;
    return true;
}

// ------------------------------------------------------------------------------------------ object transforms

// Affine maps (row-major 3x4, p -> m p + t) in the order of glam's transform_point3 / transform_vector3.
__device__ __forceinline__ f3 affine_point(const float *m, f3 p) {
    return f3{((m[0] * p.x + m[1] * p.y) + m[2] * p.z) + m[3], ((m[4] * p.x + m[5] * p.y) + m[6] * p.z) + m[7],
              ((m[8] * p.x + m[9] * p.y) + m[10] * p.z) + m[11]};
}
__device__ __forceinline__ f3 affine_vector(const float *m, f3 v) {
    return f3{(m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[4] * v.x + m[5] * v.y) + m[6] * v.z,
              (m[8] * v.x + m[9] * v.y) + m[10] * v.z};
}

// ObjectTransform::incoming_ray (object/transform.rs): the world ray mapped to object space, its direction
// normalised again (Ray::new).
__device__ __forceinline__ DRay object_ray(const GXform *x, const DRay &r) {
    return make_ray(affine_point(x->inv, r.o), affine_vector(x->inv, r.d));
}

// ------------------------------------------------------------------------------------------ SDFs

__device__ __forceinline__ float ipow(float x, int n) {  // binary powering in compiler-rt's __powisf2 order
    float r = 1.0f, a = x;
    while (n > 0) {
        if (n & 1) r = r * a;
        n >>= 1;
        if (n) a = a * a;
    }
    return r;
}

// x^(n-1) and x^n for the Mandelbulb estimate, each as ipow computes it.  For the usual power 8 the products are
// written out (ipow's 1 * a is exact): x^7 = (x * x^2) * x^4 and x^8 = x^4 * x^4, sharing x^2 and x^4.
__device__ __forceinline__ void bulb_powers(float x, int n, float *x_nm1, float *x_n) {
    if (n == 8) {
        const float x2 = x * x, x4 = x2 * x2;
        *x_nm1 = (x * x2) * x4;
        *x_n = x4 * x4;
    } else {
        *x_nm1 = ipow(x, n - 1);
        *x_n = ipow(x, n);
    }
}

// A complex number, for the Mandelbulb's angle multiples: (cos a + i sin a)^n = cos na + i sin na.
struct cx {
    float re, im;
};

// The Mandelbulb's powers of a unit complex number (|a| = 1 within rounding): a square as (2 re^2 - 1, 2 re im),
// since re^2 - im^2 = 2 re^2 - 1 on the unit circle (three instructions instead of five), a product with fused
// multiply-adds; binary powering (the lowest factor taken as is), so for n = 8 three squarings.  IEEE operations
// only (an fma rounds once), so the oracle's single precision evaluates it bit for bit.
__device__ __forceinline__ cx csq_unit(cx a) {
    const float t = a.re + a.re;
    return cx{fmaf(t, a.re, -1.0f), t * a.im};
}
__device__ __forceinline__ cx cmul_fma(cx a, cx b) {
    return cx{fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re)};
}
__device__ __forceinline__ cx cpow_unit(cx a, int n) {
    if (n == 8) return csq_unit(csq_unit(csq_unit(a)));
    cx r{1.0f, 0.0f};
    bool have = false;
    while (n > 0) {
        if (n & 1) {
            r = have ? cmul_fma(r, a) : a;
            have = true;
        }
        n >>= 1;
        if (n) a = csq_unit(a);
    }
    return r;
}

// 1/sqrt(s) by three Newton steps y <- y (3/2 - (s/2) y^2), each as y * fma(-(h y), y, 3/2) with h = s/2, from the
// bit-level first guess 0x5f3759df - (bits >> 1) (Lomont, "Fast Inverse Square Root", 2003): the guess is within
// 3.5%, the steps reach fp32 rounding (1.8e-3, 4.7e-6, then below 1e-7).  Finite for every finite s >= 0.  IEEE
// multiplies and fmas, so the oracle's single precision evaluates it bit for bit.
__device__ __forceinline__ float inv_sqrt_nr(float s) {
    float y = __uint_as_float(0x5f3759dfu - (__float_as_uint(s) >> 1));
    const float h = 0.5f * s;
    y = y * fmaf(-(h * y), y, 1.5f);
    y = y * fmaf(-(h * y), y, 1.5f);
    y = y * fmaf(-(h * y), y, 1.5f);
    return y;
}

// One iteration of the Mandelbulb estimate: r = |z|; unless r exceeds the bail-out radius, dr <- n r^(n-1) dr + 1
// and z <- z^n + p0, the spherical power (radius r^n, polar angle n theta, azimuth n phi) written without
// trigonometry: with rho = |(z.x, z.y)|, u = (z.z + i rho) / r = e^{i theta} and w = (z.x + i z.y) / rho =
// e^{i phi}, z^n = r^n (Im u^n Re w^n, Im u^n Im w^n, Re u^n); w is 1 on the z axis and u is 1 at r = 0, where
// the other factor vanishes.  Returns false when r exceeds the bail-out radius (z and dr unchanged).
// |z| and rho are rsqrt products (s * (1/sqrt s)) and the unit vectors use those reciprocals: two Newton
// reciprocal square roots instead of two correctly rounded square roots and two divisions per iteration.  The
// sums of products (|z|^2, dr, z^n + p0) are fused multiply-adds and u^n, w^n unit-circle powers.
// The SDF's bail-out radius and power come in registers (loaded once per estimate, not per iteration); POW is the
// power when the caller knows it (8, the usual one) and 0 for the run-time `power`.
template <int POW>
__device__ __forceinline__ bool bulb_iterate(float bailout, int power, f3 p0, f3 &z, float &dr, float &r) {
    const int n = POW ? POW : power;
    const float rho2 = fmaf(z.x, z.x, z.y * z.y);
    const float s2 = fmaf(z.z, z.z, rho2);  // |z|^2
    const float ir = inv_sqrt_nr(s2);
    r = s2 * ir;
    if (r > bailout) return false;
    float rp1, zr;
    bulb_powers(r, n, &rp1, &zr);
    dr = fmaf(rp1 * (float)n, dr, 1.0f);
    const float iq = inv_sqrt_nr(rho2);
    const float rho = rho2 * iq;
    cx u{1.0f, 0.0f}, w{1.0f, 0.0f};
    if (r > 0.0f) u = cx{z.z * ir, rho * ir};
    if (rho > 0.0f) w = cx{z.x * iq, z.y * iq};
    const cx un = cpow_unit(u, n), wn = cpow_unit(w, n);
    z = f3{fmaf(un.im * wn.re, zr, p0.x), fmaf(un.im * wn.im, zr, p0.y), fmaf(un.re, zr, p0.z)};
    return true;
}

// The estimate 0.5 ln(r) r / dr; at the centre itself (r = 0) its limit, 0.
__device__ __forceinline__ float bulb_estimate(float r, float dr) {
    return r == 0.0f ? 0.0f : 0.5f * dm::log_f(r) * r / dr;
}

// The smooth minimum of n sphere distances in list order: d <- min(d, d_i) - h^2 k / 4 with
// h = max(k - |d - d_i|, 0) / k, the division a multiply by 1/k and h^2 (k / 4) fused into the subtraction.
__device__ __forceinline__ float smin_spheres(const GSphere *s, int n, float k, f3 p) {
    float d = length_fma(p - f3{s[0].cx, s[0].cy, s[0].cz}) - s[0].radius;
    const float ik = 1.0f / k, k4 = k * 0.25f;
    for (int i = 1; i < n; ++i) {
        const GSphere q = s[i];
        const float di = length_fma(p - f3{q.cx, q.cy, q.cz}) - q.radius;
        const float h = fmaxf(k - fabsf(d - di), 0.0f) * ik;
        d = fmaf(-(h * h), k4, fminf(d, di));
    }
    return d;
}

__device__ float sdf_eval(const GSphere *spheres, const GSdf &f, f3 p) {
    if (f.kind == RAYNA_SDF_SPHERE) {
        const GSphere s = spheres[f.first_sphere];
        return length_fma(p - f3{s.cx, s.cy, s.cz}) - s.radius;
    }
    if (f.kind == RAYNA_SDF_SMIN_SPHERES) {
        // March items are SDF-major, so the lanes of a wave normally evaluate one SDF: its spheres are then read at
        // wave-uniform addresses and the loop is counted in scalar registers.
        const int32_t first = __builtin_amdgcn_readfirstlane(f.first_sphere);
        const int32_t n = __builtin_amdgcn_readfirstlane(f.n_spheres);
        if (__all(f.first_sphere == first && f.n_spheres == n)) return smin_spheres(spheres + first, n, f.k, p);
        return smin_spheres(spheres + f.first_sphere, f.n_spheres, f.k, p);
    }
    // Mandelbulb distance estimate
    const f3 p0 = p - f3{f.cx, f.cy, f.cz};
    f3 z = p0;
    float dr = 1.0f, r = 0.0f;
    const int iterations = f.bulb_iterations, power = f.power;
    const float bailout = f.bailout;
    if (__all(power == 8)) {
        for (int i = 0; i < iterations; ++i)
            if (!bulb_iterate<8>(bailout, power, p0, z, dr, r)) break;
    } else {
        for (int i = 0; i < iterations; ++i)
            if (!bulb_iterate<0>(bailout, power, p0, z, dr, r)) break;
    }
    return bulb_estimate(r, dr);
}

__device__ f3 sdf_normal(const SceneView &sv, const GSdf &f, f3 p) {
    const float e = f.epsilon;
    const f3 g{sdf_eval(sv.spheres, f, f3{p.x + e, p.y, p.z}) - sdf_eval(sv.spheres, f, f3{p.x - e, p.y, p.z}),
               sdf_eval(sv.spheres, f, f3{p.x, p.y + e, p.z}) - sdf_eval(sv.spheres, f, f3{p.x, p.y - e, p.z}),
               sdf_eval(sv.spheres, f, f3{p.x, p.y, p.z + e}) - sdf_eval(sv.spheres, f, f3{p.x, p.y, p.z - e})};
    return normalize(g);
}

// ------------------------------------------------------------------------------------------ closest hit

struct Best {
    float t;
    int32_t rank;
    uint32_t ref;
    uint32_t aux;
    uint32_t sub;  // an instance's hit: its primitive's reference (hit_sub)
};

__device__ __forceinline__ uint32_t make_ref(int kind, int32_t index) {
    return ((uint32_t)kind << HIT_KIND_SHIFT) | ((uint32_t)index & HIT_INDEX_MASK);
}

// defined with the generate kernel below; k_extend_packet computes the camera rays at depth 0
__device__ __forceinline__ DRay camera_ray(const GViewport &vp, uint32_t width, uint32_t height, uint32_t sdim,
                                           dm::Key key, uint32_t pixel, uint32_t k);

#include "traverse.hpp"
#include "march.hpp"

// ------------------------------------------------------------------------------------------ shading

struct SHit {
    f3 pos, normal, ray_normal;
    bool front;
    int32_t material;
};

__device__ __forceinline__ f3 unit_sphere(float r1, float r2) {  // rng::normal_on_unit_sphere (rng.rs:75-87)
    float sn, cs;
    dm::sincos_turns(r1, &sn, &cs);
    const float q = sqrtf(r2 * (1.0f - r2));
    return f3{(cs * 2.0f) * q, (sn * 2.0f) * q, 1.0f - 2.0f * r2};
}

// The primitive a hit reference names: its kind and index into that kind's array (the inner primitive or volume
// of a transformed object, or the primitive `sub` names inside an instanced mesh BVH), its material, and the ray
// and distance to evaluate it with (the object-space ray and distance of a transformed object).  x: the
// transform, or null.
struct HitPrim {
    int kind;
    int32_t prim;
    int32_t material;
    DRay ray;
    float t;
    const GXform *x;
};

__device__ __forceinline__ HitPrim hit_prim(const SceneView &sv, const DRay &ray, float t, uint32_t ref,
                                            uint32_t aux, const uint32_t *sub) {
    HitPrim p{(int)(ref >> HIT_KIND_SHIFT), (int32_t)(ref & HIT_INDEX_MASK), 0, ray, t, nullptr};
    if (p.kind == ITEM_TRIANGLE) {
        p.material = sv.tris[p.prim].material;
    } else if (p.kind == ITEM_SDF) {
        p.material = sv.sdfs[p.prim].material;
    } else {
        const GItem it = sv.items[p.prim];
        p.material = it.material;
        p.prim = it.index;
        if (p.kind == ITEM_XFORM) {
            p.x = sv.xforms + it.index;
            p.kind = p.x->inner_kind;
            p.prim = p.x->inner_index;
            p.ray = object_ray(p.x, ray);
            p.t = __uint_as_float(aux);
            if (p.kind == ITEM_LINK) {  // an instance: the primitive inside it
                const uint32_t s = *sub;
                p.kind = (int)(s >> HIT_KIND_SHIFT);
                p.prim = (int32_t)(s & HIT_INDEX_MASK);
                if (p.kind != ITEM_TRIANGLE) p.prim = sv.items[p.prim].index;
            }
        }
    }
    return p;
}

// Re-derives the full intersection record of a hit (the extend kernel keeps only distance and reference).  A
// transformed object's record is derived in object space and mapped out as outgoing_intersection does
// (object/transform.rs): the position by the object-to-world map, both normals by its linear part, normalised.
// SDF: the scene may hold SDF hits whose normal was not given (without, that case is compiled out: for scenes
// without SDFs, whose shading then keeps fewer registers live).
template <bool SDF = true>
__device__ void hit_geometry(const SceneView &sv, const DRay &ray, float t, uint32_t ref, uint32_t aux,
                             const uint32_t *sub, const dm::Key &key, uint32_t pixel, uint32_t sample,
                             uint32_t bounce, SHit *h, bool have_sdf_n = false, f3 sdf_n = f3{0.0f, 0.0f, 0.0f}) {
    const HitPrim p = hit_prim(sv, ray, t, ref, aux, sub);
    const DRay &r = p.ray;
    h->pos = at(r, p.t);
    h->material = p.material;
    switch (p.kind) {
        case ITEM_TRIANGLE: {
            const GTri g = sv.tris[p.prim];
            float tt, u, v, det;
            tri_t(g, r, p.t, p.t, &tt, &u, &v, &det);
            f3 n{0.0f, 0.0f, 0.0f};
            tri_normal(sv.tri_normals[p.prim], u, v, &n);
            h->normal = n;
            h->front = signbit(det);
            h->ray_normal = n * -dm::signum(det);
            break;
        }
        case ITEM_SPHERE: {
            const GSphere s = sv.spheres[p.prim];
            const f3 local = (h->pos - f3{s.cx, s.cy, s.cz}) / s.radius;
            const bool inside = dot(r.d, local) > 0.0f;
            h->normal = local;
            h->ray_normal = inside ? neg(local) : local;
            h->front = !inside;
            break;
        }
        case ITEM_PARALLELOGRAM: {
            const GPara g = sv.paras[p.prim];
            const f3 n{g.nx, g.ny, g.nz};
            const float denom = dot(n, r.d);
            h->normal = n;
            h->front = signbit(denom);
            h->ray_normal = neg(n) * dm::signum(denom);
            break;
        }
        case ITEM_AXIS_BOX: {
            float tt, w;
            int axis = 0;
            f3 sg{0.0f, 0.0f, 0.0f};
            box_t(sv.boxes[p.prim], with_inv(r), p.t, p.t, &tt, &axis, &w, &sg);
            f3 rn{0.0f, 0.0f, 0.0f};
            if (axis == 0) rn.x = sg.x;
            if (axis == 1) rn.y = sg.y;
            if (axis == 2) rn.z = sg.z;
            h->ray_normal = rn;
            h->normal = rn * w;
            h->front = !signbit(w);
            break;
        }
        case ITEM_VOLUME: {
            float a[4];
            dm::draw4(key, pixel, sample, bounce, dm::SLOT_MEDIUM_AUX, a);
            h->normal = unit_sphere(a[0], a[1]);
            h->ray_normal = unit_sphere(a[2], a[3]);
            h->front = true;
            break;
        }
        default: {  // ITEM_SDF: computed by k_sdf_normal when given
            const f3 n = (have_sdf_n || !SDF) ? sdf_n : sdf_normal(sv, sv.sdfs[p.prim], h->pos);
            h->normal = n;
            h->ray_normal = n;
            h->front = !(aux & HIT_AUX_BACK);
            break;
        }
    }
    if (p.x) {
        h->normal = normalize(affine_vector(p.x->fwd, h->normal));
        h->ray_normal = normalize(affine_vector(p.x->fwd, h->ray_normal));
        h->pos = affine_point(p.x->fwd, h->pos);
    }
}

// Intersection::uv and ::side of a hit (shared/intersect.rs:7-43), for the debug render modes: sphere_uv
// (sphere.rs:114-121), the triangle's barycentrics (triangle.rs:104), the plane coordinates (planar/mod.rs:191),
// the box face's uv and side (axis_box.rs:121-129: side = 2 axis + (sgn > 0), `-1.0 as u32` saturating to 0),
// the volume's random uv (volumetric.rs:153) and the SDF's march iteration (raymarched.rs:106-109, carried in the
// hit flags from bit 8).  A transform leaves both as the object-space test found them.
// sphere_uv (sphere.rs:114-121) of a unit vector.
__device__ __forceinline__ void unit_sphere_uv(f3 p, float uv[2]) {
    const float PI = 3.14159274101257324f;
    const float theta = dm::acos_f(-p.y);
    const float phi = dm::atan2_f(-p.z, p.x) + PI;
    uv[0] = phi / (2.0f * PI);
    uv[1] = theta / PI;
}

// pos_l (optional): Intersection::pos_l -- the sphere's unit local point (sphere.rs:91), the barycentrics
// (1 - u - v, u, v) (triangle.rs:95-101), the position relative to the plane's origin (planar/mod.rs:176) or the
// box centre (axis_box.rs:122), else the position; a transform maps it as a point (transform.rs:120-121).
__device__ void hit_uv_side(const SceneView &sv, const DRay &world_ray, float world_t, uint32_t ref, uint32_t aux,
                            const uint32_t *sub, const dm::Key &key, uint32_t pixel, uint32_t sample,
                            uint32_t bounce, float uv[2], int *side, f3 *pos_l = nullptr) {
    const HitPrim p = hit_prim(sv, world_ray, world_t, ref, aux, sub);
    const DRay &ray = p.ray;
    const float t = p.t;
    const int32_t idx = p.prim;
    const f3 pos = at(ray, t);
    f3 pl = pos;
    uv[0] = uv[1] = 0.0f;
    *side = 0;
    switch (p.kind) {
        case ITEM_TRIANGLE: {
            float tt, det;
            tri_t(sv.tris[idx], ray, t, t, &tt, &uv[0], &uv[1], &det);
            pl = f3{(1.0f - uv[0]) - uv[1], uv[0], uv[1]};
            break;
        }
        case ITEM_SPHERE: {
            const GSphere s = sv.spheres[idx];
            const f3 local = (pos - f3{s.cx, s.cy, s.cz}) / s.radius;
            unit_sphere_uv(local, uv);
            pl = local;
            break;
        }
        case ITEM_PARALLELOGRAM: {
            float tt;
            const GPara g = sv.paras[idx];
            para_t(g, ray, t, t, &tt, uv);
            pl = pos - f3{g.px, g.py, g.pz};
            break;
        }
        case ITEM_AXIS_BOX: {
            float tt, w;
            int axis = 0;
            f3 sg{0.0f, 0.0f, 0.0f};
            const GBox g = sv.boxes[idx];
            box_t(g, with_inv(ray), t, t, &tt, &axis, &w, &sg, uv);
            const float sa = axis == 0 ? sg.x : axis == 1 ? sg.y : sg.z;
            *side = 2 * axis + (sa > 0.0f ? 1 : 0);
            pl = pos - f3{g.cx, g.cy, g.cz};
            break;
        }
        case ITEM_VOLUME: {
            float b[4];
            dm::draw4(key, pixel, sample, bounce, dm::SLOT_MEDIUM_UV, b);
            uv[0] = b[0];
            uv[1] = b[1];
            break;
        }
        default:  // ITEM_SDF
            *side = (int)(aux >> HIT_AUX_SIDE_SHIFT);
            break;
    }
    if (pos_l) *pos_l = p.x ? affine_point(p.x->fwd, pl) : pl;
}

#include "textures.hpp"

// The material's colour at a hit: its solid colour, or its texture's value (TEX: the scene has textures).
template <bool TEX>
__device__ __forceinline__ Col material_colour(const SceneView &sv, const GMaterial &m, const DRay &ray, float t,
                                               uint32_t ref, uint32_t aux, const uint32_t *sub, const dm::Key &key,
                                               uint32_t pixel, uint32_t sample, uint32_t bounce, const SHit &h) {
    if (!TEX || m.texture < 0) return Col{m.r, m.g, m.b};
    float uv[2];
    int side;
    f3 pl;
    hit_uv_side(sv, ray, t, ref, aux, sub, key, pixel, sample, bounce, uv, &side, &pl);
    return texture_value(sv, m.texture, h.pos, pl, uv);
}

// Material::scatter (material/*.rs); false for "no scatter".
__device__ __forceinline__ bool scatter(const GMaterial &m, const DRay &ray, const SHit &h, const float u[4], f3 *dir) {
    switch (m.kind) {
        case RAYNA_MAT_LAMBERTIAN: {  // lambertian.rs:29-36
            const f3 rnd = unit_sphere(u[0], u[1]) * dm::cbrt01(u[2]);
            const f3 v = h.ray_normal + rnd;
            if (!try_normalize(v, dir)) *dir = h.ray_normal;
            return true;
        }
        case RAYNA_MAT_METAL: {  // metal.rs:17-33
            const f3 refl = reflect(ray.d, h.ray_normal);
            const f3 v = refl + unit_sphere(u[0], u[1]) * m.fuzz;
            if (dot(v, h.ray_normal) > 0.0f) {
                *dir = normalize(v);
                return true;
            }
            return false;
        }
        case RAYNA_MAT_DIELECTRIC: {  // dielectric.rs:19-39, 74-79
            const float ratio = h.front ? 1.0f / m.ior : m.ior;
            const float cos_t = fminf(dot(neg(ray.d), h.ray_normal), 1.0f);
            const float sin_t = sqrtf(1.0f - cos_t * cos_t);
            const bool tir = ratio * sin_t > 1.0f;
            const float r0 = (1.0f - ratio) / (1.0f + ratio);
            const float r0s = r0 * r0;
            const float refl = r0s + (1.0f - r0s) * ipow(1.0f - cos_t, 5);
            const bool schlick = refl > u[0];
            *dir = (tir || schlick) ? reflect(ray.d, h.ray_normal) : refract(ray.d, h.ray_normal, ratio);
            return true;
        }
        case RAYNA_MAT_ISOTROPIC:  // isotropic.rs:29-31
            *dir = unit_sphere(u[0], u[1]);
            return true;
        default:
            return false;
    }
}

// Material::reflected_light applied to the path throughput (future colour) T, with the material's colour at the
// hit.
__device__ __forceinline__ Col reflected(const GMaterial &m, const DRay &ray, const SHit &h, Col albedo, Col T) {
    switch (m.kind) {
        case RAYNA_MAT_LAMBERTIAN:
        case RAYNA_MAT_METAL: return T * albedo;
        case RAYNA_MAT_DIELECTRIC: {
            if (h.front) return T;
            const float dist = length(h.pos - ray.o);
            const float tr = -m.density * dist;
            return (T * albedo) * dm::exp_f(tr);
        }
        case RAYNA_MAT_ISOTROPIC: {
            const float dist = length(h.pos - ray.o);
            const float tr = -m.density * dist;
            return (T * albedo) * dm::exp_f(tr);
        }
        default: return Col{0.0f, 0.0f, 0.0f};
    }
}

__device__ __forceinline__ Col sky(int kind, f3 d) {  // skybox/simple.rs, skybox/none.rs (HDRI: sky_image)
    if (kind == RAYNA_SKY_WHITE) return Col{1.0f, 1.0f, 1.0f};
    if (kind == RAYNA_SKY_SIMPLE) {
        const float a = 0.5f * d.y + 0.5f;
        return Col{1.0f + (0.5f - 1.0f) * a, 1.0f + (0.7f - 1.0f) * a, 1.0f + (1.0f - 1.0f) * a};
    }
    return Col{0.0f, 0.0f, 0.0f};
}

// ------------------------------------------------------------------------------------------ validation

// shared/validate.rs (13-198): the invariants the engine asserts in debug builds along every path
// (render/renderer.rs:395-561), counted per check (RAYNA_CHECK_*) instead of asserted, with the engine's
// tolerances.  The same checks, in the same arithmetic, as the oracle's fp32 restatement.
__device__ __forceinline__ bool finite3(f3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
// validate::normal3: finite and glam's is_normalized, |n.n - 1| <= 2e-4
__device__ __forceinline__ bool unit3(f3 n) { return finite3(n) && fabsf(dot(n, n) - 1.0f) <= 2e-4f; }
// approx's relative_eq with validate.rs's EPSILON = 1e-6 and RELATIVE = 1e-3
__device__ __forceinline__ bool approx_eq(float a, float b) {
    if (a == b) return true;
    if (isinf(a) || isinf(b)) return false;
    const float d = fabsf(a - b);
    if (d <= 1e-6f) return true;
    return d <= fmaxf(fabsf(a), fabsf(b)) * 1e-3f;
}
// validate::colour: every channel not NaN, not infinite, >= 0
__device__ __forceinline__ bool colour_ok(Col c) {
    auto ok = [](float v) { return !isnan(v) && !isinf(v) && v >= 0.0f; };
    return ok(c.r) && ok(c.g) && ok(c.b);
}
__device__ __forceinline__ void count_check(uint32_t *checks, int which) { atomicAdd(checks + which, 1u); }
// validate::intersection (validate.rs:160-198) of a hit of `ray` at distance t in [0.001, FLT_MAX]
__device__ __forceinline__ void validate_hit(uint32_t *checks, const DRay &ray, float t, const SHit &h,
                                             const float uv[2]) {
    if (!isfinite(t) || !(1e-3f <= t && t <= FLT_MAX)) count_check(checks, RAYNA_CHECK_HIT_DIST);
    const f3 p = at(ray, t);
    if (!finite3(h.pos) || !approx_eq(length(ray.o - h.pos), t) || !approx_eq(h.pos.x, p.x) ||
        !approx_eq(h.pos.y, p.y) || !approx_eq(h.pos.z, p.z))
        count_check(checks, RAYNA_CHECK_HIT_POS);
    if (!unit3(h.normal) || !unit3(h.ray_normal)) count_check(checks, RAYNA_CHECK_HIT_NORMAL);
    if (!(uv[0] >= 0.0f && uv[0] <= 1.0f && uv[1] >= 0.0f && uv[1] <= 1.0f)) count_check(checks, RAYNA_CHECK_HIT_UV);
}

// ------------------------------------------------------------------------------------------ kernels

// The camera ray of sample k of pixel `pixel`: render_px_msaa's stratified jitter (renderer.rs:370-389) and
// Viewport::calc_ray (camera.rs:191-213).  k_generate stores it where a kernel loads it; k_extend_packet and the
// first k_shade compute it instead of loading it (they wait on memory more than on arithmetic).
__device__ __forceinline__ DRay camera_ray(const GViewport &vp, uint32_t width, uint32_t height, uint32_t sdim,
                                           dm::Key key, uint32_t pixel, uint32_t k) {
    const uint32_t x = pixel % width, y = pixel / width;
    float u[4];
    dm::draw4(key, pixel, k, 0, dm::SLOT_CAMERA, u);
    const uint32_t s = sdim;
    const float jx = u[0] - 0.5f, jy = u[1] - 0.5f;
    float cx, cy;
    if (k < s * s) {
        const float inv = 1.0f / (float)s;
        const uint32_t ii = k % s, jj = k / s;
        cx = ((float)x + jx * inv) + (float)ii * inv;
        cy = ((float)y + jy * inv) + (float)jj * inv;
    } else {
        cx = (float)x + jx;
        cy = (float)y + jy;
    }
    const float W = (float)width, H = (float)height;
    const float uu = (cx - W / 2.0f) / H;
    const float vv = (cy - H / 2.0f) / H;
    const f3 pc{vp.pixel_center[0], vp.pixel_center[1], vp.pixel_center[2]};
    const f3 vu{vp.vu[0], vp.vu[1], vp.vu[2]}, vvv{vp.vv[0], vp.vv[1], vp.vv[2]};
    const f3 pix = (pc + vu * uu) + vvv * vv;
    float sn, cs;
    dm::sincos_turns(u[3], &sn, &cs);
    const float rr = sqrtf(u[2]);
    const float dx = rr * cs, dy = rr * sn;
    const f3 pos{vp.pos[0], vp.pos[1], vp.pos[2]};
    const f3 du{vp.du[0], vp.du[1], vp.du[2]}, dv{vp.dv[0], vp.dv[1], vp.dv[2]};
    const f3 org = (pos + du * dx) + dv * dy;
    return make_ray(org, pix - org);
}

// Camera ray of every slot of the chunk.  Throughput and radiance get no initial values: the first k_shade takes
// them as 1 and 0 (the adjoint replay reads the primal radiance the field still holds) and writes the radiance
// of every slot, as k_shade_mode does.  BATCH: each slot finds its request (the last with slot_base <= slot) and
// records it for the later kernels; the ray itself only where the chunk stores its camera rays.
template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_generate(FrameParams fp, PathBuffers pb, BatchView bv) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= fp.n_slots) return;
    uint32_t l = i;  // slot within the request
    if (BATCH) {
        uint32_t lo = 0, hi = bv.n_reqs - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (bv.reqs[mid].slot_base <= i) lo = mid;
            else hi = mid - 1;
        }
        bv.slot_req[i] = lo;
        if (!fp.rays_stored) return;
        const RequestParams &R = bv.reqs[lo];
        fp.vp = R.vp;
        fp.width = R.width;
        fp.height = R.height;
        fp.sdim = R.sdim;
        fp.key0 = R.key0;
        fp.key1 = R.key1;
        fp.s0 = R.s0;
        fp.n_spp = R.n_spp;
        l = i - R.slot_base;
    }
    const DRay r = camera_ray(fp.vp, fp.width, fp.height, fp.sdim, dm::Key{fp.key0, fp.key1}, l / fp.n_spp,
                              fp.s0 + l % fp.n_spp);
    pb.ox[i] = r.o.x;
    pb.oy[i] = r.o.y;
    pb.oz[i] = r.o.z;
    pb.dx[i] = r.d.x;
    pb.dy[i] = r.d.y;
    pb.dz[i] = r.d.z;
}
// Adds g to the gradient of material `mat` in gradient copy `copy` (float64 hardware atomics).
__device__ __forceinline__ void grad_add(const AdjointParams &ap, uint32_t copy, int32_t mat, double g0, double g1,
                                         double g2) {
    double *p = ap.grad + ((size_t)copy * ap.n_materials + (uint32_t)mat) * 3;
    unsafeAtomicAdd(p, g0);
    unsafeAtomicAdd(p + 1, g1);
    unsafeAtomicAdd(p + 2, g2);
}

// One bounce of ray_colour_recursive in forward form: L += T * Le; scatter; T = reflected(T); paths that
// scatter below the depth limit are compacted into the next queue (one atomic per wave).
//
// ADJOINT: the same bounce replayed for path replay backpropagation (Vicini, Speierer and Jakob 2021,
// algorithm 2).  The radiance field holds R, the radiance the path still gathers from this vertex on (the
// sample's primal radiance at the camera vertex); d is the sample's share of its pixel's loss gradient.  A
// light adds d*T to its emission gradient and R -= T*Le; a scatter whose attenuation is albedo times a
// colour-independent factor (lambertian, metal, isotropic, dielectric leaving the medium) adds d*R/albedo to
// its albedo gradient (nothing for a channel whose albedo is exactly 0).  The colours never change a sampling
// decision, so the replay follows the primal path exactly.
//
// Register budget: 96 VGPRs, so that five waves share each SIMD (the kernel is latency-bound).
//
// TEX: the scene has textured materials or an image sky (without, their code is compiled out).  A textured
// material's colour receives no gradient.
//
// VALIDATE (single PBR frames): the engine's debug invariants are counted into fp.checks (validate_hit and the
// checks of validate.rs): the ray shaded, its hit, the material colour, emitted radiance, the throughput after
// a scatter and, where the path ends, its radiance.
#ifndef RAYNA_SHADE_STRIP
#define RAYNA_SHADE_STRIP 16
#endif
constexpr uint32_t SHADE_STRIP = RAYNA_SHADE_STRIP;
// Waves per SIMD the register budget of k_shade allows.  WIDE (the forward shade without textures, for scenes
// without SDFs: the SDF normal code is compiled out) runs at eight: it then spills some registers, but its loads
// (queue entry, hit, triangle, material) are a chain of dependent round trips, and more waves hide more of them
// (the 1080p mesh frame's shading 8.7 -> 8.0 ms; six waves 8.3).  Everything else keeps five: the SDF frame's
// shading went from 4.8 to 6.2 ms at eight, and the adjoint, textured and validating forms spill more.
#ifndef RAYNA_SHADE_WAVES
#define RAYNA_SHADE_WAVES 5
#endif
#ifndef RAYNA_SHADE_WAVES_WIDE
#define RAYNA_SHADE_WAVES_WIDE 8
#endif
template <bool ADJOINT, bool BATCH, bool TEX, bool VALIDATE = false, bool WIDE = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(
    (WIDE && !ADJOINT && !TEX && !VALIDATE) ? RAYNA_SHADE_WAVES_WIDE : RAYNA_SHADE_WAVES))) void k_shade(
    SceneView sv, FrameParams fp, PathBuffers pb, const int32_t *queue_in, const uint32_t *count_in,
    int32_t *queue_out, uint32_t *count_out, uint32_t depth, AdjointParams ap, BatchView bv) {
    const uint32_t count = queue_in ? *count_in : fp.n_slots;
    const uint32_t lane = __lane_id();
    const uint32_t copy = ADJOINT ? blockIdx.x % ap.copies : 0u;
    // Each wave shades strips of SHADE_STRIP * 64 consecutive queue positions; the slots that go on are staged
    // in the wave's part of LDS and appended to the next queue with one atomic per strip.
    __shared__ uint32_t staged[BLOCK * SHADE_STRIP];
    uint32_t *mine = staged + (threadIdx.x / 64) * 64 * SHADE_STRIP;
    const uint32_t n_waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = wave * 64 * SHADE_STRIP; base < count; base += n_waves * 64 * SHADE_STRIP) {
      uint32_t n_kept = 0;
      for (uint32_t j = 0; j < SHADE_STRIP; ++j) {
        const uint32_t q = base + j * 64 + lane;
        bool keep = false;
        uint32_t slot = 0;
        if (q < count) {
            slot = queue_in ? (uint32_t)queue_in[q] : q;
            const uint32_t ref = pb.hit_ref[slot];
            dm::Key key{fp.key0, fp.key1};
            uint32_t pixel, sample, ray_depth = fp.ray_depth;
            int32_t skybox = sv.skybox, sky_img = sv.sky_image;
            DRay ray;
            if (BATCH) {
                const RequestParams &R = bv.reqs[bv.slot_req[slot]];
                const uint32_t l = slot - R.slot_base;
                pixel = l / R.n_spp;
                sample = R.s0 + l % R.n_spp;
                key = dm::Key{R.key0, R.key1};
                ray_depth = R.ray_depth;
                skybox = R.skybox;
                sky_img = R.sky_image;
                if (!queue_in && !fp.rays_stored) ray = camera_ray(R.vp, R.width, R.height, R.sdim, key, pixel, sample);
            } else {
                pixel = slot / fp.n_spp;
                sample = fp.s0 + slot % fp.n_spp;
                if (!queue_in) ray = camera_ray(fp.vp, fp.width, fp.height, fp.sdim, key, pixel, sample);
            }
            // The first bounce (identity queue) starts from the camera ray with throughput 1 and radiance 0 (the
            // adjoint's radiance field holds the primal radiance); k_generate stores only the ray.  A batch chunk
            // whose camera rays k_generate stored (chunks with SDFs, whose march loads them) loads it; others
            // recompute it from the request's viewport.
            Col T{1.0f, 1.0f, 1.0f}, L{0.0f, 0.0f, 0.0f};
            if (queue_in || (BATCH && fp.rays_stored))
                ray = ray_lazy(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]}, f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
            else
                ray = ray_lazy(ray.o, ray.d);
            if (queue_in) T = Col{pb.tr[slot], pb.tg[slot], pb.tb[slot]};
            if (VALIDATE && !(finite3(ray.o) && unit3(ray.d))) count_check(fp.checks, RAYNA_CHECK_RAY);
            // Forward: the radiance this bounce adds, L += add.  After the first bounce the radiance field is read
            // and written only when add is not zero in every channel: L + 0 is L (L is never -0: it starts as
            // 0 + add, and an exact cancellation rounds to +0).
            Col add{0.0f, 0.0f, 0.0f};
            if (ADJOINT) L = Col{pb.lr[slot], pb.lg[slot], pb.lb[slot]};
            if (ref == 0u) {
                if (!ADJOINT)
                    add = T * ((TEX && skybox == RAYNA_SKY_HDRI) ? sky_image(sv, sky_img, ray.d) : sky(skybox, ray.d));
                if (VALIDATE && !colour_ok(add)) count_check(fp.checks, RAYNA_CHECK_COLOUR);
            } else {
                SHit h;
                const float t = pb.hit_t[slot];
                const uint32_t aux = pb.hit_aux[slot];
                f3 sdf_n{0.0f, 0.0f, 0.0f};
                const bool sdf_hit = (ref >> HIT_KIND_SHIFT) == (uint32_t)ITEM_SDF && pb.sdf_nx;
                if (sdf_hit) sdf_n = f3{pb.sdf_nx[slot], pb.sdf_ny[slot], pb.sdf_nz[slot]};
                hit_geometry<!WIDE>(sv, ray, t, ref, aux, pb.hit_sub + slot, key, pixel, sample, depth, &h, sdf_hit,
                                    sdf_n);
                if (VALIDATE) {
                    float uv[2];
                    int side;
                    hit_uv_side(sv, ray, t, ref, aux, pb.hit_sub + slot, key, pixel, sample, depth, uv, &side);
                    validate_hit(fp.checks, ray, t, h, uv);
                }
                const GMaterial m = sv.materials[h.material];
                const Col colour =
                    material_colour<TEX>(sv, m, ray, t, ref, aux, pb.hit_sub + slot, key, pixel, sample, depth, h);
                if (VALIDATE && !colour_ok(colour)) count_check(fp.checks, RAYNA_CHECK_COLOUR);
                const bool graded = !TEX || m.texture < 0;
                double d0 = 0.0, d1 = 0.0, d2 = 0.0;
                if (ADJOINT) {
                    d0 = (double)ap.d_image[3 * pixel] * ap.inv_spp;
                    d1 = (double)ap.d_image[3 * pixel + 1] * ap.inv_spp;
                    d2 = (double)ap.d_image[3 * pixel + 2] * ap.inv_spp;
                }
                if (m.kind == RAYNA_MAT_LIGHT) {
                    const Col le = colour;
                    if (ADJOINT) {
                        if (graded) grad_add(ap, copy, h.material, d0 * T.r, d1 * T.g, d2 * T.b);
                        L = Col{L.r - T.r * le.r, L.g - T.g * le.g, L.b - T.b * le.b};
                    } else {
                        add = T * le;
                        if (VALIDATE && !colour_ok(add)) count_check(fp.checks, RAYNA_CHECK_COLOUR);
                    }
                }
                float u[4];
                dm::draw4(key, pixel, sample, depth, dm::SLOT_SCATTER, u);
                f3 dir;
                if (scatter(m, ray, h, u, &dir)) {
                    if (ADJOINT && depth < ray_depth) {
                        const bool albedo_scales = m.kind != RAYNA_MAT_DIELECTRIC || !h.front;
                        if (albedo_scales && graded)
                            grad_add(ap, copy, h.material, m.r != 0.0f ? d0 * (double)L.r / (double)m.r : 0.0,
                                     m.g != 0.0f ? d1 * (double)L.g / (double)m.g : 0.0,
                                     m.b != 0.0f ? d2 * (double)L.b / (double)m.b : 0.0);
                    }
                    T = reflected(m, ray, h, colour, T);
                    if (VALIDATE && !colour_ok(T)) count_check(fp.checks, RAYNA_CHECK_COLOUR);
                    if (depth < ray_depth) {
                        const DRay nr = make_ray(h.pos, dir);
                        pb.ox[slot] = nr.o.x;
                        pb.oy[slot] = nr.o.y;
                        pb.oz[slot] = nr.o.z;
                        pb.dx[slot] = nr.d.x;
                        pb.dy[slot] = nr.d.y;
                        pb.dz[slot] = nr.d.z;
                        pb.tr[slot] = T.r;
                        pb.tg[slot] = T.g;
                        pb.tb[slot] = T.b;
                        keep = true;
                    }
                }
            }
            const bool adds = !(add.r == 0.0f && add.g == 0.0f && add.b == 0.0f);
            if (!ADJOINT && queue_in && (adds || VALIDATE)) L = Col{pb.lr[slot], pb.lg[slot], pb.lb[slot]};
            if (!ADJOINT) L = L + add;
            if (VALIDATE && !keep && !colour_ok(L)) count_check(fp.checks, RAYNA_CHECK_SAMPLE);
            if (ADJOINT || !queue_in || adds) {
                pb.lr[slot] = L.r;
                pb.lg[slot] = L.g;
                pb.lb[slot] = L.b;
            }
        }
        const unsigned long long mask = __ballot(keep);
        if (keep) mine[n_kept + (uint32_t)__popcll(mask & below)] = slot;
        n_kept += (uint32_t)__popcll(mask);
      }
      if (n_kept != 0u) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t wbase = 0;
        if (lane == 0u) wbase = atomicAdd(count_out, n_kept);
        wbase = __shfl(wbase, 0);
        for (uint32_t i = lane; i < n_kept; i += 64) queue_out[wbase + i] = (int32_t)mine[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
}

// Sum of the gradient copies: out[i] = sum over copies of grad[copy * n + i], n = 3 * materials.
__global__ __launch_bounds__(BLOCK) void k_sum_grad(const double *grad, uint32_t n, uint32_t copies, double *out) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (uint32_t c = 0; c < copies; ++c) s += grad[(size_t)c * n + i];
    out[i] = s;
}

// Running sums, in sample order, of the n samples (slots base + p * n + k, k < n) of each pixel p of
// [p0, p0 + np) into sum (three planes of P floats; first: start from 0).  The whole block calls it.  The samples
// are staged in LDS with coalesced loads (one row of n | 1 floats per pixel and channel: an odd stride keeps the
// summing threads on distinct banks), then one thread per (pixel, channel) adds its row in order.
__device__ void resolve_pixels(const PathBuffers &pb, uint32_t base, uint32_t p0, uint32_t np, uint32_t n, uint32_t P,
                               float *sum, bool first) {
    __shared__ float stage[3 * RESOLVE_STAGE];
    const uint32_t t = threadIdx.x;
    const float *src[3] = {pb.lr + base + (size_t)p0 * n, pb.lg + base + (size_t)p0 * n, pb.lb + base + (size_t)p0 * n};
    const uint32_t w = n | 1u;
    if (w > RESOLVE_STAGE) {  // a thread per pixel, straight from HBM
        for (uint32_t px = t; px < np; px += BLOCK)
            for (int c = 0; c < 3; ++c) {
                float s = first ? 0.0f : sum[c * P + p0 + px];
                for (uint32_t k = 0; k < n; ++k) s = s + src[c][(size_t)px * n + k];
                sum[c * P + p0 + px] = s;
            }
        return;
    }
    const uint32_t m = np * n;
    for (uint32_t i = t; i < m; i += BLOCK) {
        const uint32_t px = i / n, k = i - px * n;
        const uint32_t o = px * w + k;
        stage[o] = src[0][i];
        stage[RESOLVE_STAGE + o] = src[1][i];
        stage[2 * RESOLVE_STAGE + o] = src[2][i];
    }
    __syncthreads();
    for (uint32_t u = t; u < 3 * np; u += BLOCK) {
        const uint32_t c = u / np, px = u - c * np;
        const float *row = stage + c * RESOLVE_STAGE + px * w;
        float s = first ? 0.0f : sum[c * P + p0 + px];
        for (uint32_t k = 0; k < n; ++k) s = s + row[k];
        sum[c * P + p0 + px] = s;
    }
}

// Per-pixel running sum of the chunk's samples in sample order (render_px_msaa's sum, renderer.rs:398-402).
__global__ __launch_bounds__(BLOCK) void k_resolve(FrameParams fp, PathBuffers pb, float *sum, uint32_t chunk_samples,
                                                   int first) {
    const uint32_t per = resolve_pixels_per_block(chunk_samples);
    const uint32_t p0 = blockIdx.x * per;
    const uint32_t np = fp.n_pixels - p0 < per ? fp.n_pixels - p0 : per;
    resolve_pixels(pb, 0u, p0, np, chunk_samples, fp.n_pixels, sum, first != 0);
}

// Sample mean, accumulation buffer update and output image (accum_buffer.rs:31-45).
__global__ __launch_bounds__(BLOCK) void k_finalize(FrameParams fp, const float *sum, float *acc, double frames,
                                                    float *out, int layout) {
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= fp.n_pixels) return;
    const uint32_t P = fp.n_pixels;
    const float n = (float)fp.spp;
    const float mean[3] = {sum[p] / n, sum[P + p] / n, sum[2 * P + p] / n};
    if (fp.checks && !colour_ok(Col{mean[0], mean[1], mean[2]})) count_check(fp.checks, RAYNA_CHECK_PIXEL);
    const uint32_t x = p % fp.width, y = p / fp.width;
    const size_t o = layout == RAYNA_LAYOUT_XMAJOR ? ((size_t)x * fp.height + y) * 3 : (size_t)p * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = acc[c * P + p] + mean[c];
        acc[c * P + p] = a;
        out[o + c] = (float)((double)a / frames);
    }
}

// k_resolve for the requests of a batch chunk: blockIdx.y selects the request.
__global__ __launch_bounds__(BLOCK) void k_resolve_batch(BatchView bv, PathBuffers pb, float *sum) {
    const RequestParams &R = bv.reqs[blockIdx.y];
    const uint32_t per = resolve_pixels_per_block(R.n_spp);
    const uint32_t p0 = blockIdx.x * per;
    if (p0 >= R.n_pixels) return;  // the whole block
    const uint32_t np = R.n_pixels - p0 < per ? R.n_pixels - p0 : per;
    resolve_pixels(pb, R.slot_base, p0, np, R.n_spp, R.n_pixels, sum + R.sum_offset, R.first_piece != 0);
}

// The sample mean of each finished request of a batch (a first frame: the accumulation is the mean itself),
// in the request's layout.
__global__ __launch_bounds__(BLOCK) void k_finalize_batch(BatchView bv, const uint32_t *final_reqs, const float *sum,
                                                          float *out) {
    const RequestParams &R = bv.reqs[final_reqs[blockIdx.y]];
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t P = R.n_pixels;
    if (p >= P) return;
    const float n = (float)R.spp;
    const float *su = sum + R.sum_offset;
    const uint32_t x = p % R.width, y = p / R.width;
    const size_t o = R.sum_offset + (R.layout == RAYNA_LAYOUT_XMAJOR ? ((size_t)x * R.height + y) * 3 : (size_t)p * 3);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = 0.0f + su[c * P + p] / n;
        out[o + c] = (float)((double)a / 1.0);
    }
}

// render_px_once for the non-PBR modes (renderer.rs:428-490): the camera ray's closest hit coloured by the mode,
// the sky on a miss.
__global__ __launch_bounds__(BLOCK) void k_shade_mode(SceneView sv, FrameParams fp, PathBuffers pb) {
    const dm::Key key{fp.key0, fp.key1};
    for (uint32_t slot = blockIdx.x * BLOCK + threadIdx.x; slot < fp.n_slots; slot += gridDim.x * BLOCK) {
        const DRay ray = ray_lazy(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]}, f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
        const uint32_t ref = pb.hit_ref[slot];
        Col L;
        if (ref == 0u) {
            L = sv.skybox == RAYNA_SKY_HDRI ? sky_image(sv, sv.sky_image, ray.d) : sky(sv.skybox, ray.d);
        } else {
            const uint32_t pixel = slot / fp.n_spp, sample = fp.s0 + slot % fp.n_spp;
            const float t = pb.hit_t[slot];
            const uint32_t aux = pb.hit_aux[slot];
            SHit h;
            f3 sdf_n{0.0f, 0.0f, 0.0f};
            const bool sdf_hit = (ref >> HIT_KIND_SHIFT) == (uint32_t)ITEM_SDF && pb.sdf_nx;
            if (sdf_hit) sdf_n = f3{pb.sdf_nx[slot], pb.sdf_ny[slot], pb.sdf_nz[slot]};
            hit_geometry(sv, ray, t, ref, aux, pb.hit_sub + slot, key, pixel, sample, 0, &h, sdf_hit, sdf_n);
            const float PAL[13][3] = {{1.0f, 1.0f, 1.0f}, {1.0f, 0.0f, 0.0f}, {1.0f, 0.5f, 0.0f}, {1.0f, 1.0f, 0.0f},
                                      {0.5f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.5f}, {0.0f, 1.0f, 1.0f},
                                      {0.0f, 0.5f, 1.0f}, {0.0f, 0.0f, 1.0f}, {0.5f, 0.0f, 1.0f}, {1.0f, 0.0f, 1.0f},
                                      {0.0f, 0.0f, 0.0f}};
            f3 vis{0.0f, 0.0f, 0.0f};
            bool use_vis = true;
            switch (fp.mode) {
                case RAYNA_MODE_RAY_NORMAL: vis = h.ray_normal; break;
                case RAYNA_MODE_OUTWARD_NORMAL: vis = h.normal; break;
                case RAYNA_MODE_SCATTER: {
                    float u[4];
                    dm::draw4(key, pixel, sample, 0, dm::SLOT_SCATTER, u);
                    if (!scatter(sv.materials[h.material], ray, h, u, &vis)) vis = f3{0.0f, 0.0f, 0.0f};
                    break;
                }
                default: use_vis = false; break;
            }
            if (use_vis) {  // (f / 2.) as Channel + 0.5
                L = Col{vis.x / 2.0f + 0.5f, vis.y / 2.0f + 0.5f, vis.z / 2.0f + 0.5f};
            } else if (fp.mode == RAYNA_MODE_FRONT_FACE) {
                const int c = h.front ? 1 : 0;
                L = Col{PAL[c][0], PAL[c][1], PAL[c][2]};
            } else if (fp.mode == RAYNA_MODE_DISTANCE) {
                const float val = 2.0f * dm::cbrt01(t);
                const float fl = fminf(fmaxf(floorf(val), 0.0f), 12.0f);
                const float ce = fminf(fmaxf(ceilf(val), 0.0f), 12.0f);
                const float frac = val - fl;
                const int ia = (int)fl, ib = (int)ce;
                // Colour::lerp: a + (b - a) * frac
                L = Col{PAL[ia][0] + (PAL[ib][0] - PAL[ia][0]) * frac, PAL[ia][1] + (PAL[ib][1] - PAL[ia][1]) * frac,
                        PAL[ia][2] + (PAL[ib][2] - PAL[ia][2]) * frac};
            } else {
                float uv[2];
                int side;
                hit_uv_side(sv, ray, t, ref, aux, pb.hit_sub + slot, key, pixel, sample, 0, uv, &side);
                if (fp.mode == RAYNA_MODE_UV) {
                    L = Col{fminf(fmaxf(uv[0], 0.0f), 1.0f), fminf(fmaxf(uv[1], 0.0f), 1.0f), 0.0f};
                } else {  // RAYNA_MODE_SIDE
                    const int c = side % 12 + 1;
                    L = Col{PAL[c][0], PAL[c][1], PAL[c][2]};
                }
            }
        }
        pb.lr[slot] = L.r;
        pb.lg[slot] = L.g;
        pb.lb[slot] = L.b;
    }
}

// The UI's 8-bit conversion (rayna_ui/src/ext/img_ext.rs:15-45): c^(1/2.2) * 255 with Rust's saturating u8 cast.
__device__ __forceinline__ uint8_t channel_u8(float c) {
    const float g = dm::pow_f(c, 1.0f / 2.2f) * 255.0f;
    if (!(g > 0.0f)) return 0;
    if (g >= 255.0f) return 255;
    return (uint8_t)g;
}

__global__ __launch_bounds__(BLOCK) void k_to_srgb8(const float *img, uint32_t n_pixels, uchar4 *rgba) {
    const uint32_t p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n_pixels) return;
    rgba[p] = make_uchar4(channel_u8(img[3 * p]), channel_u8(img[3 * p + 1]), channel_u8(img[3 * p + 2]), 255);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t value, uint32_t n) {
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) p[i] = value;
}

// VALU issue-rate calibration: eight independent multiply-add chains per lane, each link one v_mul_f32 and one
// v_add_f32 (no packed or fused forms are generated), so a lane issues 16 fp32 VALU instructions per iteration
// and enough waves fill every SIMD that the dependency latency is hidden.
constexpr uint32_t VALU_PEAK_CHAINS = 8;
__global__ __launch_bounds__(BLOCK) void k_valu_peak(float *out, uint32_t iters, float m, float c) {
    const uint32_t gid = blockIdx.x * BLOCK + threadIdx.x;
    float a[VALU_PEAK_CHAINS];
#pragma unroll
    for (uint32_t k = 0; k < VALU_PEAK_CHAINS; ++k) a[k] = (float)(gid + k) * 1e-6f;
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (uint32_t k = 0; k < VALU_PEAK_CHAINS; ++k) a[k] = a[k] * m + c;
    }
    float sum = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < VALU_PEAK_CHAINS; ++k) sum += a[k];
    out[gid] = sum;
}

// Counts the bit patterns [first, first + count) on which sqrt_rn or sqrt_rn_wave differs from sqrtf (NaNs compare
// equal).  A wave holds 64 consecutive patterns, so waves whose inputs are all at least 2^-96 check the unscaled
// path of sqrt_rn_wave and the others its scaled one.
__global__ __launch_bounds__(BLOCK) void k_sqrt_check(uint32_t first, uint32_t count, unsigned long long *mismatches) {
    uint32_t bad = 0;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < count; i += gridDim.x * BLOCK) {
        const float x = __uint_as_float(first + i);
        const float a = sqrt_rn(x), c = sqrt_rn_wave(x), b = sqrtf(x);
        if (__float_as_uint(a) != __float_as_uint(b) && !(a != a && b != b)) ++bad;
        else if (__float_as_uint(c) != __float_as_uint(b) && !(c != c && b != b)) ++bad;
    }
    if (bad) atomicAdd(mismatches, (unsigned long long)bad);
}

__global__ void k_count_segments(const uint32_t *queue_count, uint32_t n, uint32_t first_count,
                                 unsigned long long *segments) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned long long s = first_count;
        for (uint32_t i = 1; i < n; ++i) s += queue_count[i];
        *segments += s;
    }
}

inline unsigned grid_for(uint64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

using ExtendKernel = void (*)(SceneView, FrameParams, PathBuffers, const int32_t *, const uint32_t *, uint32_t,
                              int2 *, uint32_t *, ExtendTuning, BatchView);
// The k_extend instantiations, [batch][instances][form]: form 0 neither spheres nor media, 1 spheres, 2 media
// (with the sphere tests, which sphere-bounded media need).
const ExtendKernel EXTEND_KERNELS[2][2][3] = {
    {{k_extend<false, false, false, false>, k_extend<false, false, true, false>, k_extend<false, false, true, true>},
     {k_extend<false, true, false, false>, k_extend<false, true, true, false>, k_extend<false, true, true, true>}},
    {{k_extend<true, false, false, false>, k_extend<true, false, true, false>, k_extend<true, false, true, true>},
     {k_extend<true, true, false, false>, k_extend<true, true, true, false>, k_extend<true, true, true, true>}}};

int extend_form(const SceneView &sv) { return sv.media ? 2 : (sv.sphere_tests ? 1 : 0); }

// Resident 128-lane blocks per CU of each instantiation, taken once per process.
struct ExtendOccupancy {
    int n[2][2][3];
};
const ExtendOccupancy &extend_occupancy() {
    static ExtendOccupancy o;
    static std::once_flag once;
    std::call_once(once, [] {
        for (int b = 0; b < 2; ++b)
            for (int i = 0; i < 2; ++i)
                for (int f = 0; f < 3; ++f) {
                    int n = 0;
                    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, EXTEND_KERNELS[b][i][f], EXT_BLOCK, 0) !=
                            hipSuccess ||
                        n < 1)
                        n = 4;
                    o.n[b][i][f] = n;
                }
    });
    return o;
}

int extend_lds_stack() { return EXT_LDS_STACK; }

int extend_blocks_per_cu() {
    const ExtendOccupancy &o = extend_occupancy();
    int n = 1;
    for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 2; ++i)
            for (int f = 0; f < 3; ++f) n = std::max(n, o.n[b][i][f]);
    return n;
}
int extend_block() { return EXT_BLOCK; }

int march_blocks_per_cu() {
    int n = 0, nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_march<false>, MARCH_BLOCK, 0) != hipSuccess || n < 1) n = 4;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_march<true>, MARCH_BLOCK, 0) == hipSuccess && nb >= 1)
        n = std::min(n, nb);
    return n;
}
int march_block() { return MARCH_BLOCK; }

void launch_generate(const FrameParams &fp, const PathBuffers &pb, const BatchView &bv, hipStream_t s) {
    if (bv.reqs)
        hipLaunchKernelGGL(k_generate<true>, dim3(grid_for(fp.n_slots, BLOCK)), dim3(BLOCK), 0, s, fp, pb, bv);
    else
        hipLaunchKernelGGL(k_generate<false>, dim3(grid_for(fp.n_slots, BLOCK)), dim3(BLOCK), 0, s, fp, pb, bv);
}

bool extend_packets(const SceneView &sv, const FrameParams &fp, const ExtendTuning &tune, const BatchView &bv,
                    uint32_t depth) {
    // a pixel's samples in neighbouring slots: wave packets
    return depth < tune.packet_depths && !sv.instances && sv.packet_need <= PACKET_STACK &&
           (bv.reqs || fp.n_spp >= PACKET_MIN_SPP);
}

void launch_extend(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue,
                   const uint32_t *count, uint32_t depth, int grid, int2 *overflow, uint32_t *work,
                   const ExtendTuning &tune, const BatchView &bv, hipStream_t s) {
    if (extend_packets(sv, fp, tune, bv, depth)) {
        const int f = extend_form(sv);
        auto packet = bv.reqs ? (f == 2   ? k_extend_packet<true, true, true>
                                 : f == 1 ? k_extend_packet<true, true, false>
                                          : k_extend_packet<true, false, false>)
                              : (f == 2   ? k_extend_packet<false, true, true>
                                 : f == 1 ? k_extend_packet<false, true, false>
                                          : k_extend_packet<false, false, false>);
        hipLaunchKernelGGL(packet, dim3(grid_for(fp.n_slots, EXT_BLOCK)), dim3(EXT_BLOCK), 0, s, sv, fp, pb, queue,
                           count, depth, bv);
        return;
    }
    const int b = bv.reqs ? 1 : 0, i = sv.instances ? 1 : 0, f = extend_form(sv);
    // `grid` is the CU count times the largest residency of any form, which the overflow columns are sized
    // for; each form launches as many blocks as stay resident (never more than `grid`)
    const int g = std::max(1, std::min(grid, std::max(1, grid / extend_blocks_per_cu()) * extend_occupancy().n[b][i][f]));
    hipLaunchKernelGGL(EXTEND_KERNELS[b][i][f], dim3(g), dim3(EXT_BLOCK), 0, s, sv, fp, pb, queue, count, depth,
                       overflow, work, tune, bv);
}

void launch_shade(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue_in,
                  const uint32_t *count_in, int32_t *queue_out, uint32_t *count_out, uint32_t depth, int grid,
                  const AdjointParams *adjoint, const BatchView &bv, hipStream_t s) {
    const bool tex = sv.textured != 0;
    const bool validate = fp.checks && !bv.reqs && !adjoint;
    const bool wide = sv.n_sdfs == 0;
    auto kernel = bv.reqs    ? (tex ? k_shade<false, true, true>
                                    : (wide ? k_shade<false, true, false, false, true> : k_shade<false, true, false>))
                  : adjoint  ? (tex ? k_shade<true, false, true> : k_shade<true, false, false>)
                  : validate ? (tex ? k_shade<false, false, true, true> : k_shade<false, false, false, true>)
                             : (tex ? k_shade<false, false, true>
                                    : (wide ? k_shade<false, false, false, false, true> : k_shade<false, false, false>));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(BLOCK), 0, s, sv, fp, pb, queue_in, count_in, queue_out, count_out,
                       depth, adjoint ? *adjoint : AdjointParams{}, bv);
}

void launch_march(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue,
                  const uint32_t *count, uint32_t n_sdf, float *mt, uint64_t mt_stride, uint32_t *work, int grid,
                  int merge_grid, const ExtendTuning &tune, const BatchView &bv, int32_t *sdf_queue,
                  uint32_t *sdf_count, hipStream_t s) {
    if (n_sdf == 0) return;
    launch_fill_u32(sdf_count, 0u, 1, s);
    if (bv.reqs) {
        hipLaunchKernelGGL(k_march<true>, dim3(grid), dim3(MARCH_BLOCK), 0, s, fp, pb, sv.sdfs, sv.spheres, queue,
                           count, n_sdf, mt, mt_stride, work, tune, bv);
        hipLaunchKernelGGL(k_march_merge<true>, dim3(merge_grid), dim3(BLOCK), 0, s, sv, fp, pb, queue, count,
                           n_sdf, mt, mt_stride, sdf_queue, sdf_count, bv);
    } else {
        hipLaunchKernelGGL(k_march<false>, dim3(grid), dim3(MARCH_BLOCK), 0, s, fp, pb, sv.sdfs, sv.spheres, queue,
                           count, n_sdf, mt, mt_stride, work, tune, bv);
        hipLaunchKernelGGL(k_march_merge<false>, dim3(merge_grid), dim3(BLOCK), 0, s, sv, fp, pb, queue, count,
                           n_sdf, mt, mt_stride, sdf_queue, sdf_count, bv);
    }
    hipLaunchKernelGGL(k_sdf_normal, dim3(merge_grid), dim3(BLOCK), 0, s, sv, pb, sdf_queue, sdf_count);
}

void launch_resolve_batch(const BatchView &bv, uint32_t blocks, const PathBuffers &pb, float *sum, hipStream_t s) {
    hipLaunchKernelGGL(k_resolve_batch, dim3(blocks, bv.n_reqs), dim3(BLOCK), 0, s, bv, pb, sum);
}

void launch_finalize_batch(const BatchView &bv, const uint32_t *final_reqs, uint32_t n_final, uint32_t max_pixels,
                           const float *sum, float *out, hipStream_t s) {
    if (n_final == 0) return;
    hipLaunchKernelGGL(k_finalize_batch, dim3(grid_for(max_pixels, BLOCK), n_final), dim3(BLOCK), 0, s, bv,
                       final_reqs, sum, out);
}

void launch_sum_grad(const double *grad, uint32_t n, uint32_t copies, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_grad, dim3(grid_for(n, BLOCK)), dim3(BLOCK), 0, s, grad, n, copies, out);
}

void launch_shade_mode(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_shade_mode, dim3(grid), dim3(BLOCK), 0, s, sv, fp, pb);
}

void launch_to_srgb8(const float *img, uint32_t n_pixels, uint8_t *rgba, hipStream_t s) {
    hipLaunchKernelGGL(k_to_srgb8, dim3(grid_for(n_pixels, BLOCK)), dim3(BLOCK), 0, s, img, n_pixels,
                       reinterpret_cast<uchar4 *>(rgba));
}

void launch_resolve(const FrameParams &fp, const PathBuffers &pb, float *sum, uint32_t chunk_samples, int first,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_resolve, dim3(grid_for(fp.n_pixels, (int)resolve_pixels_per_block(chunk_samples))),
                       dim3(BLOCK), 0, s, fp, pb, sum, chunk_samples, first);
}

void launch_finalize(const FrameParams &fp, const float *sum, float *acc, double frames, float *out, int layout,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(grid_for(fp.n_pixels, BLOCK)), dim3(BLOCK), 0, s, fp, sum, acc, frames, out,
                       layout);
}

void launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, BLOCK)), dim3(BLOCK), 0, s, p, value, n);
}

void launch_sqrt_check(uint32_t first, uint32_t count, unsigned long long *mismatches, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_sqrt_check, dim3(blocks), dim3(BLOCK), 0, s, first, count, mismatches);
}

uint64_t launch_valu_peak(float *out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    hipLaunchKernelGGL(k_valu_peak, dim3(blocks), dim3(BLOCK), 0, s, out, iters, 0.999f, 1e-3f);
    return (uint64_t)blocks * (BLOCK / 64) * iters * (2 * VALU_PEAK_CHAINS);
}

void launch_count_segments(const uint32_t *queue_count, uint32_t n, uint32_t first_count,
                           unsigned long long *segments, hipStream_t s) {
    hipLaunchKernelGGL(k_count_segments, dim3(1), dim3(64), 0, s, queue_count, n, first_count, segments);
}

// ------------------------------------------------------------------------------------------ static workspace

// The gradient copies of rayna_adjoint_static live in this file-scope array instead of memory the library
// allocates, so that entry point allocates nothing.  Every GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, not device memory: the device address of
// each GPU's copy comes from hipGetSymbolAddress, taken once per device with that device current and kept.
constexpr size_t ADJOINT_SCRATCH_DOUBLES = size_t(1) << 19;  // 4 MiB
static __device__ __align__(256) double scratch[ADJOINT_SCRATCH_DOUBLES];

size_t adjoint_scratch_doubles() { return ADJOINT_SCRATCH_DOUBLES; }

hipError_t adjoint_scratch(int device, double **ptr) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double *addr[MAX_DEVICES] = {};
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double *>(p);
    }
    *ptr = addr[device];
    return hipSuccess;
}

}  // namespace rayna

#ifdef RAYNA_TRAVERSAL_STATS
// Measurement builds only: the traversal counters summed since the last call (node pops, node expansions,
// primitive tests, wave-loop iterations per lane), then reset.
extern "C" int rayna_traversal_stats(unsigned long long out[4]) {
    if (hipDeviceSynchronize() != hipSuccess) return -1;
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(rayna::g_trav_stats), sizeof(unsigned long long) * 4) != hipSuccess)
        return -1;
    const unsigned long long zero[4] = {0, 0, 0, 0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(rayna::g_trav_stats), zero, sizeof(zero)) != hipSuccess) return -1;
    return 0;
}
#endif
