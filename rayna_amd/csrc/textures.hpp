// Textures and the image skybox (included by kernels.hip inside its anonymous namespace, after the hit
// helpers).  Each function restates the engine's in single precision with its order of operations.

// Image::get_bilinear (core/image.rs:86-110): corner coordinates floor and ceil clamped to the image, the fraction
// taken from the clamped floor; two lerps along x, one along y.  Colour::lerp is a + (b - a) * t
// (shared/math.rs:9-12).  A NaN coordinate reads texel 0, as Rust's saturating cast does.
__device__ __forceinline__ int clamp_index(float f, uint32_t n) {
    const float c = fminf(fmaxf(f, 0.0f), (float)(n - 1u));
    return c == c ? (int)c : 0;
}

__device__ __forceinline__ Col image_bilinear(const SceneView &sv, int32_t image, float px, float py) {
    const GImage im = sv.images[image];
    const float fx = floorf(px), fy = floorf(py);
    const int x1 = clamp_index(fx, im.width), x2 = clamp_index(ceilf(px), im.width);
    const int y1 = clamp_index(fy, im.height), y2 = clamp_index(ceilf(py), im.height);
    const float xl = px - fminf(fmaxf(fx, 0.0f), (float)(im.width - 1u));
    const float yl = py - fminf(fmaxf(fy, 0.0f), (float)(im.height - 1u));
    auto texel = [&](int x, int y) {
        const float *t = sv.texels + 3 * ((size_t)im.texel + (size_t)y * im.width + (size_t)x);
        return Col{t[0], t[1], t[2]};
    };
    auto lerp = [](Col a, Col b, float t) {
        return Col{a.r + (b.r - a.r) * t, a.g + (b.g - a.g) * t, a.b + (b.b - a.b) * t};
    };
    const Col c11 = texel(x1, y1), c12 = texel(x1, y2), c21 = texel(x2, y1), c22 = texel(x2, y2);
    return lerp(lerp(c11, c21, xl), lerp(c12, c22, xl), yl);
}

// HdrImageSkybox::sky_colour (skybox/hdri.rs:17-27).
__device__ __forceinline__ Col sky_image(const SceneView &sv, int32_t image, f3 d) {
    float uv[2];
    unit_sphere_uv(d, uv);
    const GImage im = sv.images[image];
    return image_bilinear(sv, image, uv[0] * (float)im.width, (1.0f - uv[1]) * (float)im.height);
}

// Improved Perlin noise (K. Perlin, "Improving Noise", SIGGRAPH 2002): quintic fade, gradients from the low four
// bits of the hashed corner, p a 512-entry permutation table.
__device__ __forceinline__ int noise_lattice(float f) {  // floor(f) mod 256, 0 for a non-finite f
    return isfinite(f) ? (int)(f - 256.0f * floorf(f / 256.0f)) : 0;
}
__device__ __forceinline__ float noise_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float noise_lerp(float t, float a, float b) { return a + t * (b - a); }
__device__ __forceinline__ float noise_grad(int hash, float x, float y, float z) {
    const int h = hash & 15;
    const float u = h < 8 ? x : y;
    const float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
__device__ float perlin(const int32_t *p, float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int X = noise_lattice(fx), Y = noise_lattice(fy), Z = noise_lattice(fz);
    x -= fx;
    y -= fy;
    z -= fz;
    const float u = noise_fade(x), v = noise_fade(y), w = noise_fade(z);
    const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z, B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return noise_lerp(w,
                      noise_lerp(v, noise_lerp(u, noise_grad(p[AA], x, y, z), noise_grad(p[BA], x - 1.0f, y, z)),
                                 noise_lerp(u, noise_grad(p[AB], x, y - 1.0f, z),
                                            noise_grad(p[BB], x - 1.0f, y - 1.0f, z))),
                      noise_lerp(v,
                                 noise_lerp(u, noise_grad(p[AA + 1], x, y, z - 1.0f),
                                            noise_grad(p[BA + 1], x - 1.0f, y, z - 1.0f)),
                                 noise_lerp(u, noise_grad(p[AB + 1], x, y - 1.0f, z - 1.0f),
                                            noise_grad(p[BB + 1], x - 1.0f, y - 1.0f, z - 1.0f))));
}

// ColourSource::get (texture/noise.rs:36-45) of ScalePoint(noise).set_scale(scale): the noise (one table, or one
// per channel) at the scaled point, mapped from [-1, 1] to [0, 1].
__device__ __forceinline__ Col noise_colour(const SceneView &sv, const GTexture &g, f3 q) {
    q = q * g.scale;
    const float n0 = perlin(sv.perms + g.perm[0], q.x, q.y, q.z);
    float n1 = n0, n2 = n0;
    if (g.rgb) {
        n1 = perlin(sv.perms + g.perm[1], q.x, q.y, q.z);
        n2 = perlin(sv.perms + g.perm[2], q.x, q.y, q.z);
    }
    return Col{n0 / 2.0f + 0.5f, n1 / 2.0f + 0.5f, n2 / 2.0f + 0.5f};
}

// do_checker (texture/checker.rs:44-63): the floors summed from 0; even when rem_euclid(sum, 2) < 1.
__device__ __forceinline__ bool checker_even(float sum) {
    float r = fmodf(sum, 2.0f);
    if (r < 0.0f) r = r + 2.0f;
    return r < 1.0f;
}

// Texture::value of texture `t` at a hit with world position pw, local position pl and uv.  Checkers name earlier
// textures, so the walk ends.
__device__ Col texture_value(const SceneView &sv, int32_t t, f3 pw, f3 pl, const float uv[2]) {
    for (;;) {
        const GTexture g = sv.textures[t];
        switch (g.kind) {
            case RAYNA_TEX_SOLID: return Col{g.r, g.g, g.b};
            case RAYNA_TEX_WORLD_CHECKER: {
                const float sum = ((0.0f + floorf(pw.x / g.scale + g.ox)) + floorf(pw.y / g.scale + g.oy)) +
                                  floorf(pw.z / g.scale + g.oz);
                t = checker_even(sum) ? g.even : g.odd;
                break;
            }
            case RAYNA_TEX_UV_CHECKER: {
                const float sum = (0.0f + floorf(uv[0] / g.scale + g.ox)) + floorf(uv[1] / g.scale + g.oy);
                t = checker_even(sum) ? g.even : g.odd;
                break;
            }
            case RAYNA_TEX_IMAGE: {  // ImageTexture::value (texture/image.rs:29-39)
                const GImage im = sv.images[g.image];
                const float u = g.ox + uv[0] * g.sx, v = 1.0f - (g.oy + uv[1] * g.sy);
                return image_bilinear(sv, g.image, u * (float)im.width, v * (float)im.height);
            }
            case RAYNA_TEX_UV_NOISE: return noise_colour(sv, g, f3{uv[0], uv[1], 0.0f});
            case RAYNA_TEX_WORLD_NOISE: return noise_colour(sv, g, pw);
            default: return noise_colour(sv, g, pl);  // RAYNA_TEX_LOCAL_NOISE
        }
    }
}
