// oracle/twin_capi.cpp — C entry points of the CPU twin, loaded with ctypes by the tests and the benchmark.
//
// TEST INFRASTRUCTURE ONLY (the oracle).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load liboracle.so; the product never does.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "twin.hpp"

using namespace twin;

extern "C" {

struct twin_scene {
    Scene s;
};

// builder: 0 = sweep restatement, 1 = literal O(n^2) restatement of calculate_optimal_split.
twin_scene *twin_scene_create(const rayna_scene_desc *d, int builder) {
    twin_scene *t = new twin_scene();
    build_scene(t->s, *d, builder == 1);
    return t;
}

void twin_scene_destroy(twin_scene *t) { delete t; }

uint64_t twin_scene_tree_hash(const twin_scene *t) { return t->s.tree_hash; }

// Depth-first order of the bounded primitives: object and local index per rank.
int32_t twin_scene_dfs_order(const twin_scene *t, int32_t *obj, int32_t *local, int32_t cap) {
    const int n = (int)t->s.rank_object.size();
    for (int i = 0; i < n && i < cap; ++i) {
        obj[i] = t->s.rank_object[i];
        local[i] = t->s.rank_local[i];
    }
    return n;
}

// Node count and children list of one BVH (-1 = TLAS, else the BLAS of mesh `mesh`), depth-first:
// for every node, item (or -1) and child count.  Used to compare with the product's tree.
int32_t twin_scene_tree_walk(const twin_scene *t, int32_t mesh, int32_t *items, int32_t *nchild, int32_t cap) {
    const Bvh *b = nullptr;
    if (mesh < 0) b = &t->s.tlas;
    else if (mesh < (int)t->s.mesh_blas.size() && t->s.mesh_blas[mesh] >= 0) b = &t->s.blas[t->s.mesh_blas[mesh]];
    if (!b || b->root < 0) return 0;
    int n = 0;
    std::vector<int> stack{b->root};
    while (!stack.empty()) {
        const int node = stack.back();
        stack.pop_back();
        if (n < cap) {
            items[n] = b->nodes[node].item;
            nchild[n] = (int32_t)b->nodes[node].children.size();
        }
        ++n;
        const auto &ch = b->nodes[node].children;
        for (int c = (int)ch.size() - 1; c >= 0; --c) stack.push_back(ch[c]);
    }
    return n;
}

// 0 if the camera yields a viewport, 1 if calculate_viewport fails (render_failed path).
int32_t twin_camera_invalid(const rayna_camera *cam) { return calculate_viewport(*cam).valid ? 0 : 1; }

// One frame: per-pixel sample means of rows row_start, row_start + row_step, ... and of every col_step-th
// column (row-major, 3 floats per pixel).  precision 32 or 64; box_mode 0 = engine boxes, 1 = tight conservative boxes; integrator
// 0 = recursive, 1 = forward.  Returns the number of closest-hit queries; *seconds gets the wall time.
uint64_t twin_render(const twin_scene *t, const rayna_camera *cam, const rayna_render_opts *opts, int32_t precision,
                     int32_t box_mode, int32_t integrator, uint32_t frame, uint32_t row_start, uint32_t row_step,
                     uint32_t col_step, int32_t threads, float *out, double *seconds, uint64_t *prim_tests) {
    RenderArgs a{&t->s, cam, opts, frame, box_mode, integrator, row_start, row_step == 0 ? 1u : row_step, threads,
                 col_step == 0 ? 1u : col_step};
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t segs;
    if (precision == 32) segs = render_rows<float>(a, out, prim_tests);
    else segs = render_rows<double>(a, out, prim_tests);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return segs;
}

// twin_render with validate.rs's invariants counted along the forward integrator's paths (precision 32 or 64):
// checks gets twin::N_CHECKS counters (twin::CHECK_*).
uint64_t twin_render_checked(const twin_scene *t, const rayna_camera *cam, const rayna_render_opts *opts,
                             int32_t precision, int32_t box_mode, uint32_t frame, int32_t threads, float *out,
                             uint64_t *checks) {
    RenderArgs a{&t->s, cam, opts, frame, box_mode, INTEGRATOR_FORWARD, 0, 1, threads, 1, checks};
    return precision == 32 ? render_rows<float>(a, out, nullptr) : render_rows<double>(a, out, nullptr);
}

// The UI's 8-bit conversion (twin::to_srgb8) of an x-major width x height x 3 float image to RGBA, row-major.
void twin_to_srgb8(const float *img, uint32_t w, uint32_t h, uint8_t *rgba) { to_srgb8(img, w, h, rgba); }

// Replaces the colour (albedo, or emission for a light) of the first n materials: 3 floats each.
void twin_scene_set_colours(twin_scene *t, const float *rgb, int32_t n) {
    for (int32_t i = 0; i < n && i < (int32_t)t->s.materials.size(); ++i)
        for (int c = 0; c < 3; ++c) t->s.materials[i].colour[c] = rgb[3 * i + c];
}

// Path-replay adjoint of one frame (twin::adjoint_frame): d_image is height x width x 3 floats, row-major;
// grad (3 doubles per material) is overwritten.  Returns the replay's closest-hit queries.
uint64_t twin_adjoint(const twin_scene *t, const rayna_camera *cam, const rayna_render_opts *opts, int32_t precision,
                      int32_t box_mode, uint32_t frame, int32_t threads, const float *d_image, double *grad,
                      double *seconds) {
    RenderArgs a{&t->s, cam, opts, frame, box_mode, INTEGRATOR_FORWARD, 0, 1, threads, 1};
    std::fill(grad, grad + 3 * t->s.materials.size(), 0.0);
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t segs = precision == 32 ? adjoint_frame<float>(a, d_image, grad)
                                          : adjoint_frame<double>(a, d_image, grad);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return segs;
}

// Closest hit of the camera ray of sample `k` of every pixel (row-major): depth-first rank (-1 miss,
// -2 - i for unbounded object i) and distance.
void twin_primary_hits(const twin_scene *t, const rayna_camera *cam, const rayna_render_opts *opts, int32_t precision,
                       int32_t box_mode, uint32_t k, int32_t *rank, double *dist) {
    const Viewport vp = calculate_viewport(*cam);
    const RngKey key = make_key(opts->seed, 0);
    for (uint32_t y = 0; y < opts->height; ++y)
        for (uint32_t x = 0; x < opts->width; ++x) {
            const size_t i = (size_t)y * opts->width + x;
            if (precision == 32) {
                const Ray<float> ray = camera_ray<float>(vp, *opts, key, x, y, k);
                Ctx<float> c{&t->s, key, (uint32_t)i, k, 0, box_mode, nullptr};
                Hit<float> h;
                const Interval<float> iv{true, true, 1e-3f, std::numeric_limits<float>::max()};
                if (scene_intersect(c, ray, iv, h)) {
                    rank[i] = h.rank;
                    dist[i] = h.t;
                } else {
                    rank[i] = -1;
                    dist[i] = 0;
                }
            } else {
                const Ray<double> ray = camera_ray<double>(vp, *opts, key, x, y, k);
                Ctx<double> c{&t->s, key, (uint32_t)i, k, 0, box_mode, nullptr};
                Hit<double> h;
                const Interval<double> iv{true, true, 1e-3, std::numeric_limits<double>::max()};
                if (scene_intersect(c, ray, iv, h)) {
                    rank[i] = h.rank;
                    dist[i] = h.t;
                } else {
                    rank[i] = -1;
                    dist[i] = 0;
                }
            }
        }
}

// Closest hit of one ray (origin, direction; the direction is normalised as Ray::new does).
// out: t, pos[3], normal[3], ray_normal[3], front_face; returns rank (-1 on a miss).
int32_t twin_trace_ray(const twin_scene *t, int32_t precision, int32_t box_mode, const double o[3], const double d[3],
                       double out[11], int32_t *material) {
    const RngKey key = make_key(0, 0);
    auto fill = [&](auto &h) {
        out[0] = h.t;
        out[1] = h.pos.x;
        out[2] = h.pos.y;
        out[3] = h.pos.z;
        out[4] = h.normal.x;
        out[5] = h.normal.y;
        out[6] = h.normal.z;
        out[7] = h.ray_normal.x;
        out[8] = h.ray_normal.y;
        out[9] = h.ray_normal.z;
        out[10] = h.front_face ? 1 : 0;
        *material = h.material;
        return h.rank;
    };
    if (precision == 32) {
        const Ray<float> ray = Ray<float>::make(cast3<float>(o), cast3<float>(d));
        Ctx<float> c{&t->s, key, 0, 0, 0, box_mode, nullptr};
        Hit<float> h;
        const Interval<float> iv{true, true, 1e-3f, std::numeric_limits<float>::max()};
        if (!scene_intersect(c, ray, iv, h)) return -1;
        return fill(h);
    }
    const Ray<double> ray = Ray<double>::make(cast3<double>(o), cast3<double>(d));
    Ctx<double> c{&t->s, key, 0, 0, 0, box_mode, nullptr};
    Hit<double> h;
    const Interval<double> iv{true, true, 1e-3, std::numeric_limits<double>::max()};
    if (!scene_intersect(c, ray, iv, h)) return -1;
    return fill(h);
}

// ---- probes of the arithmetic building blocks ----

void twin_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    Philox4x32::generate(ctr, key, out);
}

void twin_math_f32(int32_t fn, const float *x, const float *y, float *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        float s, c;
        switch (fn) {
            case 0: sincos_turns_f32(x[i], &s, &c); out[i] = s; break;
            case 1: sincos_turns_f32(x[i], &s, &c); out[i] = c; break;
            case 2: out[i] = log_f32(x[i]); break;
            case 3: out[i] = exp_f32(x[i]); break;
            case 4: out[i] = cbrt01_f32(x[i]); break;
            case 5: sincos_rad_f32(x[i], &s, &c); out[i] = s; break;
            case 6: sincos_rad_f32(x[i], &s, &c); out[i] = c; break;
            case 7: out[i] = atan2_f32(x[i], y[i]); break;
            case 8: out[i] = acos_f32(x[i]); break;
            default: out[i] = 0; break;
        }
    }
}

// Samplers: 0 = unit sphere normal (r1, r2), 1 = in unit ball (r1, r2, r3), 2 = in unit disk (r1, r2).
void twin_sampler_f64(int32_t kind, const double *u, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        const double *a = u + 3 * i;
        V3<double> v{0, 0, 0};
        if (kind == 0) v = unit_sphere<double>(a[0], a[1]);
        else if (kind == 1) v = unit_sphere<double>(a[0], a[1]) * std::cbrt(a[2]);
        else {
            double sn, cs;
            sincos_turns_f64(a[1], &sn, &cs);
            const double r = std::sqrt(a[0]);
            v = V3<double>{r * cs, r * sn, 0};
        }
        out[3 * i] = v.x;
        out[3 * i + 1] = v.y;
        out[3 * i + 2] = v.z;
    }
}

// Material scatter/attenuation probe in double precision for closed-form tests.  hit: pos[3],
// normal[3], ray_normal[3], front_face; ray: pos[3], dir[3] (normalised here); u: 4 uniforms.
// out: dir[3] (normalised as Ray::new does), attenuation rgb of a white future colour.  Returns 1 if the
// material scatters.
int32_t twin_scatter_f64(const rayna_material *m, const double ray_in[6], const double hit[10], const double u[4],
                         double out[6]) {
    const Ray<double> ray = Ray<double>::make(cast3<double>(ray_in), cast3<double>(ray_in + 3));
    Hit<double> h{};
    h.pos = cast3<double>(hit);
    h.normal = cast3<double>(hit + 3);
    h.ray_normal = cast3<double>(hit + 6);
    h.front_face = hit[9] != 0;
    h.t = length(h.pos - ray.pos);
    V3<double> dir;
    const int ok = scatter<double>(*m, ray, h, u, &dir) ? 1 : 0;
    if (ok) dir = normalize(dir);
    const Col a = reflected<double>(*m, ray, h, Col{m->colour[0], m->colour[1], m->colour[2]}, Col{1, 1, 1});
    out[0] = dir.x;
    out[1] = dir.y;
    out[2] = dir.z;
    out[3] = a.r;
    out[4] = a.g;
    out[5] = a.b;
    return ok;
}

// Sky colour of a direction (normalised here).
// The value of texture `tex` (0-based) at a hit with world position pos, local position pos_l and uv.
void twin_texture_value(const twin_scene *t, int32_t tex, int32_t precision, const double pos[3],
                        const double pos_l[3], const double uv[2], float out[3]) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        Hit<R> h{};
        h.pos = cast3<R>(pos);
        h.pos_l = cast3<R>(pos_l);
        h.uv[0] = (R)uv[0];
        h.uv[1] = (R)uv[1];
        const Col c = texture_value<R>(t->s, tex, h);
        out[0] = c.r;
        out[1] = c.g;
        out[2] = c.b;
    };
    if (precision == 32) run(0.0f);
    else run(0.0);
}

// The scene's sky colour (any skybox, the image sky included) for a direction (normalised as Ray::new does).
void twin_scene_sky(const twin_scene *t, int32_t precision, const double d[3], float out[3]) {
    auto run = [&](auto zero) {
        using R = decltype(zero);
        const Ray<R> ray = Ray<R>::make(V3<R>{R(0), R(0), R(0)}, cast3<R>(d));
        const Col c = sky_colour<R>(t->s, ray);
        out[0] = c.r;
        out[1] = c.g;
        out[2] = c.b;
    };
    if (precision == 32) run(0.0f);
    else run(0.0);
}

// The permutation table of an improved Perlin noise seed (512 entries).
void twin_noise_permutation(uint32_t seed, int32_t out[512]) { noise_permutation(seed, out); }

void twin_sky(int32_t kind, const double d[3], float out[3]) {
    const Ray<double> ray = Ray<double>::make(V3<double>{0, 0, 0}, cast3<double>(d));
    const Col c = sky<double>(kind, ray);
    out[0] = c.r;
    out[1] = c.g;
    out[2] = c.b;
}

// Camera ray of a pixel sample in double precision: origin[3], dir[3].
void twin_camera_ray(const rayna_camera *cam, const rayna_render_opts *opts, uint32_t x, uint32_t y, uint32_t k,
                     double out[6]) {
    const Viewport vp = calculate_viewport(*cam);
    const RngKey key = make_key(opts->seed, 0);
    const Ray<double> r = camera_ray<double>(vp, *opts, key, x, y, k);
    out[0] = r.pos.x;
    out[1] = r.pos.y;
    out[2] = r.pos.z;
    out[3] = r.dir.x;
    out[4] = r.dir.y;
    out[5] = r.dir.z;
}

// The paths of every sample of pixel (x, y) of frame 0 (Tracer::path_ranks): for sample k, codes
// [k * cap, k * cap + n_k) and n_k in lens[k].  Used to name the cause where the precisions' radiance differs.
void twin_path_ranks(const twin_scene *t, const rayna_camera *cam, const rayna_render_opts *opts, int32_t precision,
                     int32_t box_mode, uint32_t x, uint32_t y, int32_t *codes, int32_t cap, int32_t *lens) {
    const Viewport vp = calculate_viewport(*cam);
    const RngKey key = make_key(opts->seed, 0);
    for (uint32_t k = 0; k < opts->samples; ++k) {
        if (precision == 32) {
            Tracer<float> tr{&t->s, key, y * opts->width + x, k, opts->ray_depth, box_mode};
            lens[k] = tr.path_ranks(camera_ray<float>(vp, *opts, key, x, y, k), codes + (size_t)k * cap, cap);
        } else {
            Tracer<double> tr{&t->s, key, y * opts->width + x, k, opts->ray_depth, box_mode};
            lens[k] = tr.path_ranks(camera_ray<double>(vp, *opts, key, x, y, k), codes + (size_t)k * cap, cap);
        }
    }
}

// SDF value at points (n x 3) in the given precision.
void twin_sdf_eval(const twin_scene *t, int32_t sdf, int32_t precision, const double *p, int64_t n, double *out) {
    const rayna_sdf &f = t->s.sdfs[sdf];
    for (int64_t i = 0; i < n; ++i) {
        if (precision == 32) out[i] = sdf_eval<float>(t->s, f, cast3<float>(p + 3 * i));
        else out[i] = sdf_eval<double>(t->s, f, cast3<double>(p + 3 * i));
    }
}

}  // extern "C"
