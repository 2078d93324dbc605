// The C ABI of include/rayna_gpu.h: a drop-in for rayna_engine's Renderer (render/renderer.rs:36-264) that runs
// the per-pixel path tracing as wavefront HIP kernels on one gfx950 GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rayna_gpu.h"
#include "host_util.hpp"
#include "kernels.hpp"
#include "scene_build.hpp"

using namespace rayna;

namespace {

std::string g_create_error;

// rayna_debug_fail_launches: < 0 off, else the number of launch checks still to pass before one fails
std::atomic<int32_t> g_fail_launches{-1};

// Paths in flight per chunk: 2^27 (a 1080p frame of 64 samples in one chunk; ~10 GB of path state, sized for
// the 288 GB of HBM of one MI355X).  Whole-frame chunks keep every depth's launch full: at 2^23 paths the
// per-depth tails of 16 chunks cost the 1080p frame 40% of its time.
constexpr uint64_t TARGET_SLOTS = 1ull << 27;

struct DevScene {
    std::vector<void *> allocs;
    SceneView view{};
    int2 *overflow = nullptr;
};

}  // namespace

struct rayna_renderer {
    int device = 0;
    hipStream_t stream = nullptr;
    PackedScene packed;
    DevScene dev;
    rayna_camera cam{};
    rayna_render_opts opts{};
    bool cam_valid = false;
    GViewport vp{};

    // path state
    uint64_t slot_capacity = 0;
    uint32_t chunk_samples = 1;
    std::vector<void *> path_allocs;
    PathBuffers pb{};
    int32_t *queue[2] = {nullptr, nullptr};
    uint32_t *queue_count = nullptr;
    uint32_t *claims = nullptr;  // extend claim counters, EXTEND_CLAIM_WORDS per depth and one set for queries
    unsigned long long *segments = nullptr;
    float *sum = nullptr, *acc = nullptr, *out_dev = nullptr;
    uint8_t *rgba_dev = nullptr;  // the 8-bit output of rayna_render_srgb8
    uint64_t rgba_cap = 0;
    uint32_t alloc_pixels = 0;
    uint32_t queue_count_len = 0;

    uint64_t frames = 0;  // AccumulationBuffer::frame_count (accum_buffer.rs:77)
    bool acc_dirty = true;
    // frame whose per-sample radiance the radiance field still holds (one-chunk frames only), else -1
    int64_t primal_frame = -1;

    // adjoint
    float *dimg = nullptr;        // dloss/d(frame mean), row-major pixels x 3
    double *grad = nullptr;       // grad_copies x n_materials x 3
    double *grad_sum = nullptr;   // n_materials x 3
    uint32_t grad_materials = 0, grad_copies = 0;
    int grid_extend = 1024, grid_shade = 1024, grid_march = 1024;
    int cus = 1;
    float *mt = nullptr;  // unbounded-object march outcomes, slot_capacity * n_sdfs
    void *sdf_buf = nullptr;  // scenes with unbounded objects: SDF-hit normals (three planes), their queue, a counter
    uint64_t sdf_cap = 0;
    int32_t *sdf_queue = nullptr;
    uint32_t *sdf_count = nullptr;
    uint64_t mt_cap = 0;
    uint64_t chunk_target = TARGET_SLOTS;
    ExtendTuning tune;
    std::vector<hipEvent_t> events;
    // debug validation (shared/validate.rs): RAYNA_CHECKS counters, counted in PBR frames while `validate`
    uint32_t *checks = nullptr;
    // non-empty once a launch on this renderer failed (the failure's text): every later call is refused
    std::string lost;
    bool validate = false;
    std::string error;
};

namespace {

rayna_status fail(rayna_renderer *r, rayna_status st, const std::string &msg) {
    if (r) r->error = msg;
    else g_create_error = msg;
    return st;
}

#define HIP_TRY(r, expr)                                                                                      \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return fail((r), e_ == hipErrorOutOfMemory ? RAYNA_OUT_OF_MEMORY : RAYNA_DEVICE_ERROR,           \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                                   \
    } while (0)

// HIP_TRY for the calls that wait on a renderer's launches: a failure there leaves the renderer lost.
#define HIP_TRY_LAUNCHED(r, expr)                                                                             \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess) {                                                                               \
            (r)->lost = std::string(#expr) + ": " + hipGetErrorString(e_);                                    \
            return fail((r), RAYNA_DEVICE_ERROR, (r)->lost);                                                  \
        }                                                                                                     \
    } while (0)

// A lost renderer's stream and buffers are in an unknown state: it answers every call with
// RAYNA_RENDERER_LOST until the caller destroys it and creates a new one.
#define REQUIRE_LIVE(r)                                                                                       \
    do {                                                                                                      \
        if (!(r)->lost.empty())                                                                               \
            return fail((r), RAYNA_RENDERER_LOST,                                                             \
                        "renderer lost after \"" + (r)->lost + "\": destroy it and create a new one");        \
    } while (0)

// The status of the launches just issued on r's stream (hipGetLastError), or the launch failure that
// rayna_debug_fail_launches planted for this check.  Either failure leaves r lost.
rayna_status launch_status(rayna_renderer *r, const char *where) {
    hipError_t e = hipGetLastError();
    int32_t n = g_fail_launches.load();
    while (n >= 0 && !g_fail_launches.compare_exchange_weak(n, n - 1)) {
    }
    const bool injected = n == 0;
    if (injected) e = hipErrorLaunchFailure;
    if (e == hipSuccess) return RAYNA_OK;
    r->lost = std::string(where) + ": " + hipGetErrorString(e) + (injected ? " (injected)" : "");
    return fail(r, RAYNA_DEVICE_ERROR, r->lost);
}

template <class T>
rayna_status upload(rayna_renderer *r, DevScene &ds, const std::vector<T> &v, const T **out) {
    *out = nullptr;
    if (v.empty()) return RAYNA_OK;
    void *p = nullptr;
    HIP_TRY(r, hipMalloc(&p, v.size() * sizeof(T)));
    ds.allocs.push_back(p);
    HIP_TRY(r, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(p);
    return RAYNA_OK;
}

void free_scene(DevScene &ds) {
    for (void *p : ds.allocs) (void)hipFree(p);
    ds.allocs.clear();
    ds.view = SceneView{};
    ds.overflow = nullptr;
}

rayna_status load_scene(rayna_renderer *r, const rayna_scene_desc *d) {
    PackedScene ps;
    std::string err;
    const rayna_status st = pack_scene(*d, ps, err);
    if (st != RAYNA_OK) return fail(r, st, err);
    if (ps.stack_need > 4096) return fail(r, RAYNA_UNSUPPORTED, "scene BVH is deeper than 4096 stack entries");
    DevScene ds;
    SceneView &v = ds.view;
    rayna_status s2;
#define UP(vec, field) \
    if ((s2 = upload(r, ds, ps.vec, &v.field)) != RAYNA_OK) { free_scene(ds); return s2; }
    ps.nodes = device_nodes(ps.nodes);
    UP(nodes, nodes)
    UP(tris, tris)
    UP(tri_normals, tri_normals)
    UP(items, items)
    UP(spheres, spheres)
    UP(paras, paras)
    UP(boxes, boxes)
    UP(volumes, volumes)
    UP(sdfs, sdfs)
    UP(materials, materials)
    UP(xforms, xforms)
    UP(textures, textures)
    UP(images, images)
    UP(texels, texels)
    UP(perms, perms)
#undef UP
    v.root = ps.root;
    v.n_sdfs = (int32_t)ps.sdfs.size();
    v.skybox = ps.skybox;
    v.stack_need = ps.stack_need;
    v.packet_need = ps.packet_need;
    v.instances = has_instances(ps) ? 1 : 0;
    v.sphere_tests = has_spheres(ps) ? 1 : 0;
    v.media = ps.volumes.empty() ? 0 : 1;
    v.textured = has_textures(ps) ? 1 : 0;
    v.sky_image = ps.sky_image;
    // traversal stack overflow beyond the LDS entries, one column per extend thread
    const size_t deep = (size_t)std::max(0, ps.stack_need - extend_lds_stack());
    if (deep > 0) {
        void *p = nullptr;
        const size_t bytes = deep * (size_t)r->grid_extend * extend_block() * sizeof(int2);
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            free_scene(ds);
            return fail(r, e == hipErrorOutOfMemory ? RAYNA_OUT_OF_MEMORY : RAYNA_DEVICE_ERROR,
                        std::string("overflow stack: ") + hipGetErrorString(e));
        }
        ds.allocs.push_back(p);
        ds.overflow = static_cast<int2 *>(p);
    }
    free_scene(r->dev);
    r->dev = ds;
    r->packed = std::move(ps);
    return RAYNA_OK;
}

void free_paths(rayna_renderer *r) {
    for (void *p : r->path_allocs) (void)hipFree(p);
    r->path_allocs.clear();
    r->pb = PathBuffers{};
    r->queue[0] = r->queue[1] = nullptr;
    r->queue_count = nullptr;
    r->claims = nullptr;
    r->segments = nullptr;
    r->sum = r->acc = r->out_dev = nullptr;
    r->dimg = nullptr;
    r->primal_frame = -1;
    r->slot_capacity = 0;
    r->alloc_pixels = 0;
    r->queue_count_len = 0;
}

// The march outcome buffer: two planes (outcome, march iteration) of one float per (slot, unbounded object);
// and for scenes with unbounded objects the SDF-hit normals (three planes of one float per slot), the queue of
// SDF-hit slots and its counter.  Sets the path buffers' normal planes (free_paths clears them).
rayna_status ensure_mt(rayna_renderer *r) {
    const uint64_t need = r->slot_capacity * (uint64_t)std::max(0, r->dev.view.n_sdfs);
    if (need > r->mt_cap) {
        if (r->mt) (void)hipFree(r->mt);
        r->mt = nullptr;
        r->mt_cap = 0;
        HIP_TRY(r, hipMalloc((void **)&r->mt, 2 * need * sizeof(float)));
        r->mt_cap = need;
    }
    const uint64_t cap = r->slot_capacity;
    if (r->dev.view.n_sdfs > 0 && cap > r->sdf_cap) {
        if (r->sdf_buf) (void)hipFree(r->sdf_buf);
        r->sdf_buf = nullptr;
        r->sdf_cap = 0;
        HIP_TRY(r, hipMalloc(&r->sdf_buf, cap * (3 * sizeof(float) + sizeof(int32_t)) + sizeof(uint32_t)));
        r->sdf_cap = cap;
    }
    if (r->dev.view.n_sdfs > 0) {
        float *f = static_cast<float *>(r->sdf_buf);
        const uint64_t c = r->sdf_cap;
        r->pb.sdf_nx = f;
        r->pb.sdf_ny = f + c;
        r->pb.sdf_nz = f + 2 * c;
        r->sdf_queue = reinterpret_cast<int32_t *>(f + 3 * c);
        r->sdf_count = reinterpret_cast<uint32_t *>(r->sdf_queue + c);
    } else {
        r->pb.sdf_nx = r->pb.sdf_ny = r->pb.sdf_nz = nullptr;
    }
    return RAYNA_OK;
}

rayna_status alloc_paths(rayna_renderer *r) {
    const rayna_render_opts &o = r->opts;
    const uint64_t P = (uint64_t)o.width * o.height;
    // march items (slot, unbounded object) are counted in 32 bits
    const uint64_t n_sdf = (uint64_t)std::max(1, r->dev.view.n_sdfs);
    const uint64_t target = std::min<uint64_t>(r->chunk_target, 0xffffffffull / n_sdf);
    uint64_t chunk = std::max<uint64_t>(1, target / std::max<uint64_t>(1, P));
    chunk = std::min<uint64_t>(chunk, o.samples);
    const uint64_t slots = P * chunk;
    if (slots > 0x7fffffffull || slots * n_sdf > 0xffffffffull)
        return fail(r, RAYNA_UNSUPPORTED, "image too large for one chunk of paths");
    // [0, D+2) paths entering each depth; [D+2, 2(D+2)) extend ray counters; [2(D+2), 3(D+2)) march item
    // counters; [3(D+2)], [3(D+2)+1]: the primary-hit query's two counters
    const uint32_t qlen = 3 * (o.ray_depth + 2) + 2;
    if (slots <= r->slot_capacity && P <= r->alloc_pixels && qlen <= r->queue_count_len) {
        // the buffers are large enough: keep them (grow-only)
        r->chunk_samples = (uint32_t)chunk;
        r->acc_dirty = true;
        r->primal_frame = -1;
        return RAYNA_OK;
    }
    free_paths(r);
    auto alloc = [&](size_t bytes, void **out) -> rayna_status {
        HIP_TRY(r, hipMalloc(out, std::max<size_t>(bytes, 4)));
        r->path_allocs.push_back(*out);
        return RAYNA_OK;
    };
    rayna_status st;
    float **fields[] = {&r->pb.ox, &r->pb.oy, &r->pb.oz, &r->pb.dx, &r->pb.dy, &r->pb.dz, &r->pb.tr,
                        &r->pb.tg, &r->pb.tb, &r->pb.lr, &r->pb.lg, &r->pb.lb, &r->pb.hit_t};
    for (float **f : fields)
        if ((st = alloc(slots * sizeof(float), (void **)f)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&r->pb.hit_ref)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&r->pb.hit_aux)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&r->pb.hit_sub)) != RAYNA_OK) return st;
    for (int q = 0; q < 2; ++q)
        if ((st = alloc(slots * sizeof(int32_t), (void **)&r->queue[q])) != RAYNA_OK) return st;
    if ((st = alloc(qlen * sizeof(uint32_t), (void **)&r->queue_count)) != RAYNA_OK) return st;
    if ((st = alloc((size_t)(o.ray_depth + 3) * EXTEND_CLAIM_WORDS * sizeof(uint32_t), (void **)&r->claims)) !=
        RAYNA_OK)
        return st;
    if ((st = alloc(sizeof(unsigned long long), (void **)&r->segments)) != RAYNA_OK) return st;
    if ((st = alloc(3 * P * sizeof(float), (void **)&r->sum)) != RAYNA_OK) return st;
    if ((st = alloc(3 * P * sizeof(float), (void **)&r->acc)) != RAYNA_OK) return st;
    if ((st = alloc(3 * P * sizeof(float), (void **)&r->out_dev)) != RAYNA_OK) return st;
    if ((st = alloc(3 * P * sizeof(float), (void **)&r->dimg)) != RAYNA_OK) return st;
    r->pb.hit_rank = nullptr;
    r->slot_capacity = slots;
    r->chunk_samples = (uint32_t)chunk;
    r->alloc_pixels = (uint32_t)P;
    r->queue_count_len = qlen;
    r->acc_dirty = true;
    return RAYNA_OK;
}

rayna_status check_opts(rayna_renderer *r, const rayna_render_opts *o) {
    if (!o || o->width == 0 || o->height == 0 || o->samples == 0 || o->ray_branching == 0)
        return fail(r, RAYNA_INVALID_ARGUMENT, "render options must have non-zero width, height, samples, branching");
    if (o->ray_branching != 1) return fail(r, RAYNA_UNSUPPORTED, "ray_branching > 1 is not supported on the GPU");
    if (o->mode < RAYNA_MODE_PBR || o->mode > RAYNA_MODE_SIDE) return fail(r, RAYNA_INVALID_ARGUMENT, "unknown render mode");
    if (o->ray_depth > 1000) return fail(r, RAYNA_UNSUPPORTED, "ray_depth above 1000");
    return RAYNA_OK;
}

rayna_status set_grid(rayna_renderer *r) {
    hipDeviceProp_t prop;
    HIP_TRY(r, hipGetDeviceProperties(&prop, r->device));
    const int cus = std::max(1, prop.multiProcessorCount);
    r->cus = cus;
    r->grid_extend = cus * extend_blocks_per_cu();
    r->grid_march = cus * march_blocks_per_cu();
    r->grid_shade = cus * SHADE_BLOCKS_PER_CU;
    return RAYNA_OK;
}

FrameParams frame_params(const rayna_renderer *r, uint64_t frame_index) {
    FrameParams fp;
    fp.vp = r->vp;
    fp.width = r->opts.width;
    fp.height = r->opts.height;
    fp.spp = r->opts.samples;
    fp.sdim = isqrt(r->opts.samples);
    fp.key0 = (uint32_t)r->opts.seed;
    fp.key1 = (uint32_t)(r->opts.seed >> 32) ^ (uint32_t)frame_index;
    fp.ray_depth = r->opts.ray_depth;
    fp.n_pixels = r->opts.width * r->opts.height;
    fp.s0 = 0;
    fp.n_spp = 1;
    fp.n_slots = 0;
    fp.mode = r->opts.mode;
    return fp;
}

// Trace the samples [s0, s0 + n) of every pixel: generate, then extend/shade per depth.  With `adjoint`, the
// replay of path replay backpropagation: the radiance field must hold the chunk's primal radiance.
void trace_chunk(rayna_renderer *r, FrameParams &fp, uint32_t s0, uint32_t n, std::vector<hipEvent_t> *ev, size_t *evi,
                 uint64_t *extend_launches, const AdjointParams *adjoint = nullptr) {
    fp.s0 = s0;
    fp.n_spp = n;
    fp.n_slots = fp.n_pixels * n;
    const uint32_t D2 = r->opts.ray_depth + 2;
    launch_fill_u32(r->queue_count, 0u, 3 * D2, r->stream);
    launch_fill_u32(r->claims, 0u, D2 * EXTEND_CLAIM_WORDS, r->stream);
    // the camera rays are stored only for the kernels that load them: k_extend and k_march at depth 0 and
    // k_shade_mode (k_extend_packet and the first k_shade compute them)
    fp.rays_stored = fp.mode != RAYNA_MODE_PBR || r->dev.view.n_sdfs > 0 ||
                     !extend_packets(r->dev.view, fp, r->tune, BatchView{}, 0);
    if (fp.rays_stored) launch_generate(fp, r->pb, BatchView{}, r->stream);
    for (uint32_t depth = 0; depth <= r->opts.ray_depth; ++depth) {
        const int32_t *qin = depth == 0 ? nullptr : r->queue[depth & 1];
        const uint32_t *cin = depth == 0 ? nullptr : r->queue_count + depth;
        int32_t *qout = r->queue[(depth + 1) & 1];
        uint32_t *cout = r->queue_count + depth + 1;
        if (ev) (void)hipEventRecord((*ev)[(*evi)++], r->stream);
        uint32_t *work = r->claims + (size_t)depth * EXTEND_CLAIM_WORDS;
        launch_extend(r->dev.view, fp, r->pb, qin, cin, depth, r->grid_extend, r->dev.overflow, work, r->tune,
                      BatchView{}, r->stream);
        launch_march(r->dev.view, fp, r->pb, qin, cin, (uint32_t)r->dev.view.n_sdfs, r->mt, r->mt_cap,
                     r->queue_count + 2 * D2 + depth, r->grid_march, r->grid_shade, r->tune, BatchView{},
                     r->sdf_queue, r->sdf_count, r->stream);
        if (fp.mode != RAYNA_MODE_PBR) {  // render_px_once: the camera ray's hit coloured by the mode
            if (ev) (void)hipEventRecord((*ev)[(*evi)++], r->stream);
            launch_shade_mode(r->dev.view, fp, r->pb, r->grid_shade, r->stream);
            if (ev) (void)hipEventRecord((*ev)[(*evi)++], r->stream);
            ++*extend_launches;
            break;
        }
        if (ev) (void)hipEventRecord((*ev)[(*evi)++], r->stream);
        launch_shade(r->dev.view, fp, r->pb, qin, cin, qout, cout, depth, r->grid_shade, adjoint, BatchView{},
                     r->stream);
        if (ev) (void)hipEventRecord((*ev)[(*evi)++], r->stream);
        ++*extend_launches;
    }
    launch_count_segments(r->queue_count, r->opts.ray_depth + 1, fp.n_slots, r->segments, r->stream);
}

rayna_status ensure_events(rayna_renderer *r, size_t n) {
    while (r->events.size() < n) {
        hipEvent_t e;
        HIP_TRY(r, hipEventCreate(&e));
        r->events.push_back(e);
    }
    return RAYNA_OK;
}

// Device times from the events of one call: events[0] before the first kernel, events[evi - 1] after the
// last, and per trace_chunk depth the triple [before extend, after extend, after shade] in between (each pass
// over a chunk records one triple per depth, ray_depth + 1 of them; one for the non-PBR modes).
rayna_status event_times(rayna_renderer *r, size_t evi, rayna_render_stats *st) {
    float ms = 0;
    HIP_TRY(r, hipEventElapsedTime(&ms, r->events[0], r->events[evi - 1]));
    st->kernel_ms = ms;
    const size_t per_pass = r->opts.mode == RAYNA_MODE_PBR ? (size_t)r->opts.ray_depth + 1 : 1;
    double ext = 0, sh = 0, ext0 = 0;
    for (size_t i = 1, k = 0; i + 2 < evi; i += 3, ++k) {
        float a = 0, b = 0;
        HIP_TRY(r, hipEventElapsedTime(&a, r->events[i], r->events[i + 1]));
        HIP_TRY(r, hipEventElapsedTime(&b, r->events[i + 1], r->events[i + 2]));
        ext += a;
        sh += b;
        if (k % per_pass == 0) ext0 += a;
    }
    st->extend_ms = ext;
    st->shade_ms = sh;
    st->extend_camera_ms = ext0;
    return RAYNA_OK;
}

void write_checkerboard(uint32_t w, uint32_t h, float *out, int layout) {  // renderer.rs:269-292
    for (uint32_t x = 0; x < w; ++x)
        for (uint32_t y = 0; y < h; ++y) {
            const size_t o = layout == RAYNA_LAYOUT_XMAJOR ? ((size_t)x * h + y) * 3 : ((size_t)y * w + x) * 3;
            const bool black = (x + y) % 2 == 0;
            out[o] = black ? 0.0f : 1.0f;
            out[o + 1] = 0.0f;
            out[o + 2] = black ? 0.0f : 1.0f;
        }
}

uint32_t ilog2(uint32_t v) {
    uint32_t r = 0;
    while (v >>= 1) ++r;
    return r;
}

}  // namespace

extern "C" {

rayna_status rayna_create(const rayna_scene_desc *scene, const rayna_camera *camera, const rayna_render_opts *opts,
                          rayna_renderer **out) {
    g_create_error.clear();
    if (!scene || !camera || !opts || !out) return fail(nullptr, RAYNA_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    rayna_renderer *r = new rayna_renderer();
    auto bail = [&](rayna_status st) {
        g_create_error = r->error;
        rayna_destroy(r);
        return st;
    };
    rayna_status st = check_opts(r, opts);
    if (st != RAYNA_OK) return bail(st);
    if (hipGetDevice(&r->device) != hipSuccess) {
        r->error = "no HIP device";
        return bail(RAYNA_DEVICE_ERROR);
    }
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
        r->error = "hipStreamCreate failed";
        return bail(RAYNA_DEVICE_ERROR);
    }
    if ((st = set_grid(r)) != RAYNA_OK) return bail(st);
    r->opts = *opts;
    r->cam = *camera;
    r->cam_valid = viewport(r->cam, r->vp);
    if ((st = load_scene(r, scene)) != RAYNA_OK) return bail(st);
    if ((st = alloc_paths(r)) != RAYNA_OK) return bail(st);
    *out = r;
    return RAYNA_OK;
}

rayna_status rayna_set_scene(rayna_renderer *r, const rayna_scene_desc *scene) {
    if (!r || !scene) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    (void)hipSetDevice(r->device);
    rayna_status st = load_scene(r, scene);
    if (st != RAYNA_OK) return st;
    if ((st = alloc_paths(r)) != RAYNA_OK) return st;  // the chunk depends on the unbounded-object count
    r->primal_frame = -1;
    return rayna_clear_accumulation(r);
}

rayna_status rayna_set_camera(rayna_renderer *r, const rayna_camera *camera) {
    if (!r || !camera) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    r->cam = *camera;
    r->cam_valid = viewport(r->cam, r->vp);
    r->primal_frame = -1;
    return rayna_clear_accumulation(r);
}

rayna_status rayna_set_options(rayna_renderer *r, const rayna_render_opts *opts) {
    if (!r) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    rayna_status st = check_opts(r, opts);
    if (st != RAYNA_OK) return st;
    (void)hipSetDevice(r->device);
    const bool same_shape = opts->width == r->opts.width && opts->height == r->opts.height &&
                            opts->samples == r->opts.samples && opts->ray_depth == r->opts.ray_depth;
    r->opts = *opts;
    if (!same_shape && (st = alloc_paths(r)) != RAYNA_OK) return st;
    r->primal_frame = -1;
    return rayna_clear_accumulation(r);
}

rayna_status rayna_clear_accumulation(rayna_renderer *r) {
    if (!r) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    r->frames = 0;
    r->acc_dirty = true;
    return RAYNA_OK;
}

rayna_status rayna_render(rayna_renderer *r, float *out_rgb, int32_t layout, uint32_t out_dims[2],
                          rayna_render_stats *stats) {
    if (!r) return fail(r, RAYNA_INVALID_ARGUMENT, "null renderer");
    REQUIRE_LIVE(r);
    if (layout != RAYNA_LAYOUT_XMAJOR && layout != RAYNA_LAYOUT_ROWMAJOR)
        return fail(r, RAYNA_INVALID_ARGUMENT, "unknown layout");
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(r, hipSetDevice(r->device));
    {
        const rayna_status ms = ensure_mt(r);
        if (ms != RAYNA_OK) return ms;
    }
    rayna_render_stats st{};
    const uint32_t W = r->opts.width, H = r->opts.height;
    if (!r->cam_valid) {
        const uint32_t w2 = ilog2(W), h2 = ilog2(H);
        if (out_rgb) write_checkerboard(w2, h2, out_rgb, layout);
        if (out_dims) {
            out_dims[0] = w2;
            out_dims[1] = h2;
        }
        st.accum_frames = r->frames;
    } else {
        if (out_dims) {
            out_dims[0] = W;
            out_dims[1] = H;
        }
        const uint64_t P = (uint64_t)W * H;
        if (r->acc_dirty) {
            HIP_TRY(r, hipMemsetAsync(r->acc, 0, 3 * P * sizeof(float), r->stream));
            r->acc_dirty = false;
        }
        FrameParams fp = frame_params(r, r->frames);
        if (r->validate && r->opts.mode == RAYNA_MODE_PBR) fp.checks = r->checks;
        const int64_t frame_index = (int64_t)r->frames;
        r->frames += 1;
        HIP_TRY(r, hipMemsetAsync(r->segments, 0, sizeof(unsigned long long), r->stream));
        const uint32_t spp = r->opts.samples;
        const uint32_t n_chunks = (spp + r->chunk_samples - 1) / r->chunk_samples;
        std::vector<hipEvent_t> *ev = nullptr;
        const size_t n_events = (size_t)n_chunks * (r->opts.ray_depth + 1) * 3 + 2;
        if (stats) ev = &r->events;
        if (ev) {
            const rayna_status es = ensure_events(r, n_events);
            if (es != RAYNA_OK) return es;
        }
        size_t evi = 0;
        if (ev) HIP_TRY(r, hipEventRecord((*ev)[evi++], r->stream));
        uint64_t launches = 0;
        for (uint32_t c = 0; c < n_chunks; ++c) {
            const uint32_t s0 = c * r->chunk_samples;
            const uint32_t n = std::min(r->chunk_samples, spp - s0);
            trace_chunk(r, fp, s0, n, ev, &evi, &launches);
            fp.n_slots = fp.n_pixels * n;
            launch_resolve(fp, r->pb, r->sum, n, c == 0 ? 1 : 0, r->stream);
        }
        launch_finalize(fp, r->sum, r->acc, (double)r->frames, r->out_dev, layout, r->stream);
        r->primal_frame = n_chunks == 1 ? frame_index : -1;
        if (ev) HIP_TRY(r, hipEventRecord((*ev)[evi++], r->stream));
        {
            const rayna_status ls = launch_status(r, "render");
            if (ls != RAYNA_OK) return ls;
        }
        if (out_rgb)
            HIP_TRY_LAUNCHED(
                r, hipMemcpyAsync(out_rgb, r->out_dev, 3 * P * sizeof(float), hipMemcpyDeviceToHost, r->stream));
        unsigned long long segs = 0;
        HIP_TRY_LAUNCHED(r, hipMemcpyAsync(&segs, r->segments, sizeof(segs), hipMemcpyDeviceToHost, r->stream));
        HIP_TRY_LAUNCHED(r, hipStreamSynchronize(r->stream));
        st.accum_frames = r->frames;
        st.samples = P * spp;
        st.segments = segs;
        st.extend_launches = launches;
        if (ev) {
            const rayna_status ts = event_times(r, evi, &st);
            if (ts != RAYNA_OK) return ts;
        }
    }
    st.duration_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return RAYNA_OK;
}

rayna_status rayna_render_srgb8(rayna_renderer *r, uint8_t *out_rgba, uint32_t out_dims[2], rayna_render_stats *stats) {
    if (!r || !out_rgba) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    uint32_t dims[2] = {0, 0};
    const rayna_status st = rayna_render(r, nullptr, RAYNA_LAYOUT_ROWMAJOR, dims, stats);
    if (st != RAYNA_OK) return st;
    if (out_dims) {
        out_dims[0] = dims[0];
        out_dims[1] = dims[1];
    }
    const uint64_t P = (uint64_t)dims[0] * dims[1];
    if (!r->cam_valid) {  // the checkerboard: channels of exactly 0 and 1 map to 0 and 255
        std::vector<float> chk(3 * P);
        write_checkerboard(dims[0], dims[1], chk.data(), RAYNA_LAYOUT_ROWMAJOR);
        for (uint64_t p = 0; p < P; ++p) {
            for (int c = 0; c < 3; ++c) out_rgba[4 * p + c] = chk[3 * p + c] >= 1.0f ? 255 : 0;
            out_rgba[4 * p + 3] = 255;
        }
        return RAYNA_OK;
    }
    if (r->rgba_cap < 4 * P) {
        if (r->rgba_dev) (void)hipFree(r->rgba_dev);
        r->rgba_dev = nullptr;
        r->rgba_cap = 0;
        HIP_TRY(r, hipMalloc((void **)&r->rgba_dev, 4 * P));
        r->rgba_cap = 4 * P;
    }
    launch_to_srgb8(r->out_dev, (uint32_t)P, r->rgba_dev, r->stream);
    {
        const rayna_status ls = launch_status(r, "render_srgb8");
        if (ls != RAYNA_OK) return ls;
    }
    HIP_TRY_LAUNCHED(r, hipMemcpyAsync(out_rgba, r->rgba_dev, 4 * P, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY_LAUNCHED(r, hipStreamSynchronize(r->stream));
    return RAYNA_OK;
}

rayna_status rayna_set_tuning(rayna_renderer *r, const rayna_tuning *t) {
    if (!r || !t) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    if (t->refill_lanes > 64 || t->wave_batch > 4096 || t->shade_blocks_per_cu > 64)
        return fail(r, RAYNA_INVALID_ARGUMENT, "tuning out of range");
    r->tune.refill_lanes = t->refill_lanes ? t->refill_lanes : ExtendTuning{}.refill_lanes;
    r->tune.wave_batch = t->wave_batch ? t->wave_batch : ExtendTuning{}.wave_batch;
    if (t->packet_depths < -1 || t->packet_depths > 16) return fail(r, RAYNA_INVALID_ARGUMENT, "tuning out of range");
    r->tune.packet_depths = t->packet_depths == 0 ? ExtendTuning{}.packet_depths
                                                   : (uint32_t)std::max(0, t->packet_depths);
    r->grid_shade = r->cus * (int)(t->shade_blocks_per_cu ? t->shade_blocks_per_cu : SHADE_BLOCKS_PER_CU);
    const uint64_t target = t->chunk_slots ? t->chunk_slots : TARGET_SLOTS;
    if (target != r->chunk_target) {
        r->chunk_target = target;
        (void)hipSetDevice(r->device);
        return alloc_paths(r);  // clears accumulation
    }
    return RAYNA_OK;
}

rayna_status rayna_set_validate(rayna_renderer *r, int32_t enable) {
    if (!r) return fail(r, RAYNA_INVALID_ARGUMENT, "null renderer");
    REQUIRE_LIVE(r);
    HIP_TRY(r, hipSetDevice(r->device));
    if (enable && !r->checks) {
        void *p = nullptr;
        HIP_TRY(r, hipMalloc(&p, RAYNA_CHECKS * sizeof(uint32_t)));
        r->checks = static_cast<uint32_t *>(p);
    }
    if (enable) {
        launch_fill_u32(r->checks, 0u, RAYNA_CHECKS, r->stream);
        HIP_TRY(r, hipStreamSynchronize(r->stream));
    }
    r->validate = enable != 0;
    return RAYNA_OK;
}

rayna_status rayna_validate_counts(rayna_renderer *r, uint64_t *counts, int32_t reset) {
    if (!r || !counts) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    uint32_t c[RAYNA_CHECKS] = {};
    if (r->checks) {
        HIP_TRY(r, hipSetDevice(r->device));
        HIP_TRY(r, hipMemcpyAsync(c, r->checks, sizeof(c), hipMemcpyDeviceToHost, r->stream));
        if (reset) launch_fill_u32(r->checks, 0u, RAYNA_CHECKS, r->stream);
        HIP_TRY(r, hipStreamSynchronize(r->stream));
    }
    for (int i = 0; i < RAYNA_CHECKS; ++i) counts[i] = c[i];
    return RAYNA_OK;
}

rayna_status rayna_set_material_colours(rayna_renderer *r, const float *rgb, uint32_t n) {
    if (!r || (n && !rgb)) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    if (n > r->packed.materials.size()) return fail(r, RAYNA_INVALID_ARGUMENT, "more colours than materials");
    for (uint32_t i = 0; i < 3 * n; ++i)
        if (!std::isfinite(rgb[i])) return fail(r, RAYNA_INVALID_ARGUMENT, "non-finite material colour");
    HIP_TRY(r, hipSetDevice(r->device));
    for (uint32_t i = 0; i < n; ++i) {
        GMaterial &m = r->packed.materials[i];
        m.r = rgb[3 * i];
        m.g = rgb[3 * i + 1];
        m.b = rgb[3 * i + 2];
    }
    if (n) {
        // the stream may still be reading the table
        HIP_TRY(r, hipStreamSynchronize(r->stream));
        HIP_TRY(r, hipMemcpy(const_cast<GMaterial *>(r->dev.view.materials), r->packed.materials.data(),
                             n * sizeof(GMaterial), hipMemcpyHostToDevice));
    }
    r->primal_frame = -1;
    return rayna_clear_accumulation(r);
}

rayna_status rayna_adjoint(rayna_renderer *r, uint32_t frame, const float *d_image, int32_t layout, double *grad_rgb,
                           uint32_t n_materials, rayna_render_stats *stats) {
    if (!r || !d_image || !grad_rgb) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    if (layout != RAYNA_LAYOUT_XMAJOR && layout != RAYNA_LAYOUT_ROWMAJOR)
        return fail(r, RAYNA_INVALID_ARGUMENT, "unknown layout");
    if (n_materials != r->packed.materials.size())
        return fail(r, RAYNA_INVALID_ARGUMENT, "n_materials is not the scene's material count");
    if (!r->cam_valid) return fail(r, RAYNA_INVALID_ARGUMENT, "camera is invalid (the frame is the checkerboard)");
    if (r->opts.mode != RAYNA_MODE_PBR) return fail(r, RAYNA_UNSUPPORTED, "the adjoint is of RenderMode::PBR frames");
    if (n_materials == 0) return RAYNA_OK;
    const auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(r, hipSetDevice(r->device));
    {
        const rayna_status ms = ensure_mt(r);
        if (ms != RAYNA_OK) return ms;
    }
    const uint32_t W = r->opts.width, H = r->opts.height;
    const uint64_t P = (uint64_t)W * H;
    // gradient copies: up to 64, within 64 MiB
    const uint64_t per_copy = (uint64_t)n_materials * 3 * sizeof(double);
    const uint32_t copies = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, (64ull << 20) / per_copy));
    if (r->grad_materials != n_materials || r->grad_copies != copies) {
        if (r->grad) (void)hipFree(r->grad);
        if (r->grad_sum) (void)hipFree(r->grad_sum);
        r->grad = r->grad_sum = nullptr;
        r->grad_materials = r->grad_copies = 0;
        HIP_TRY(r, hipMalloc((void **)&r->grad, per_copy * copies));
        HIP_TRY(r, hipMalloc((void **)&r->grad_sum, per_copy));
        r->grad_materials = n_materials;
        r->grad_copies = copies;
    }
    // d_image in row-major pixel order
    std::vector<float> d(3 * P);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const size_t src = layout == RAYNA_LAYOUT_XMAJOR ? ((size_t)x * H + y) * 3 : ((size_t)y * W + x) * 3;
            const size_t dst = ((size_t)y * W + x) * 3;
            d[dst] = d_image[src];
            d[dst + 1] = d_image[src + 1];
            d[dst + 2] = d_image[src + 2];
        }
    HIP_TRY(r, hipMemcpyAsync(r->dimg, d.data(), d.size() * sizeof(float), hipMemcpyHostToDevice, r->stream));
    HIP_TRY(r, hipMemsetAsync(r->grad, 0, per_copy * copies, r->stream));
    HIP_TRY(r, hipMemsetAsync(r->segments, 0, sizeof(unsigned long long), r->stream));

    const uint32_t spp = r->opts.samples;
    const uint32_t n_chunks = (spp + r->chunk_samples - 1) / r->chunk_samples;
    const bool reuse = n_chunks == 1 && r->primal_frame == (int64_t)frame;
    FrameParams fp = frame_params(r, frame);
    AdjointParams ap{r->dimg, 1.0 / (double)spp, r->grad, n_materials, copies};
    const size_t n_events = (size_t)n_chunks * (r->opts.ray_depth + 1) * 3 * 2 + 2;
    if (stats) {
        const rayna_status es = ensure_events(r, n_events);
        if (es != RAYNA_OK) return es;
    }
    std::vector<hipEvent_t> *ev = stats ? &r->events : nullptr;
    size_t evi = 0;
    if (ev) HIP_TRY(r, hipEventRecord((*ev)[evi++], r->stream));
    uint64_t launches = 0;
    for (uint32_t c = 0; c < n_chunks; ++c) {
        const uint32_t s0 = c * r->chunk_samples;
        const uint32_t n = std::min(r->chunk_samples, spp - s0);
        if (!reuse) trace_chunk(r, fp, s0, n, ev, &evi, &launches);  // primal pass: the chunk's radiance
        trace_chunk(r, fp, s0, n, ev, &evi, &launches, &ap);         // adjoint replay
    }
    launch_sum_grad(r->grad, 3 * n_materials, copies, r->grad_sum, r->stream);
    if (ev) HIP_TRY(r, hipEventRecord((*ev)[evi++], r->stream));
    {
        const rayna_status ls = launch_status(r, "adjoint");
        if (ls != RAYNA_OK) return ls;
    }
    HIP_TRY_LAUNCHED(r, hipMemcpyAsync(grad_rgb, r->grad_sum, per_copy, hipMemcpyDeviceToHost, r->stream));
    unsigned long long segs = 0;
    HIP_TRY_LAUNCHED(r, hipMemcpyAsync(&segs, r->segments, sizeof(segs), hipMemcpyDeviceToHost, r->stream));
    HIP_TRY_LAUNCHED(r, hipStreamSynchronize(r->stream));
    // the radiance field now holds what the replay left of it
    r->primal_frame = -1;
    if (stats) {
        rayna_render_stats st{};
        const rayna_status ts = event_times(r, evi, &st);
        if (ts != RAYNA_OK) return ts;
        st.accum_frames = r->frames;
        st.samples = P * spp;
        st.segments = segs;
        st.extend_launches = launches;
        st.duration_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *stats = st;
    }
    return RAYNA_OK;
}

rayna_status rayna_adjoint_static(rayna_renderer *r, uint32_t frame, const float *d_image_dev, double *grad_dev,
                                  uint32_t n_materials) {
    if (!r || !d_image_dev || !grad_dev) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    if (n_materials != r->packed.materials.size())
        return fail(r, RAYNA_INVALID_ARGUMENT, "n_materials is not the scene's material count");
    if (!r->cam_valid) return fail(r, RAYNA_INVALID_ARGUMENT, "camera is invalid (the frame is the checkerboard)");
    if (r->opts.mode != RAYNA_MODE_PBR) return fail(r, RAYNA_UNSUPPORTED, "the adjoint is of RenderMode::PBR frames");
    if (n_materials == 0) return RAYNA_OK;
    const size_t cap = adjoint_scratch_doubles();
    const uint64_t per_copy = 3ull * n_materials;  // doubles
    if (per_copy > cap)
        return fail(r, RAYNA_INVALID_ARGUMENT,
                    "the static workspace holds the gradients of at most " + std::to_string(cap / 3) + " materials");
    // the handle's march buffers must already be in place (rayna_render or rayna_adjoint sets them up)
    const uint64_t mt_need = r->slot_capacity * (uint64_t)std::max(0, r->dev.view.n_sdfs);
    if (mt_need > r->mt_cap || (r->dev.view.n_sdfs > 0 && r->slot_capacity > r->sdf_cap))
        return fail(r, RAYNA_INVALID_ARGUMENT, "render a frame first: the renderer's march buffers are not in place");
    HIP_TRY(r, hipSetDevice(r->device));
    double *ws = nullptr;
    HIP_TRY(r, adjoint_scratch(r->device, &ws));
    const uint32_t copies = (uint32_t)std::min<uint64_t>(64, cap / per_copy);
    // the workspace keeps the last call's contents: clear the copies (a fill kernel, no memset) and the counter
    launch_fill_u32(reinterpret_cast<uint32_t *>(ws), 0u, (uint32_t)(2 * per_copy * copies), r->stream);
    launch_fill_u32(reinterpret_cast<uint32_t *>(r->segments), 0u, 2, r->stream);
    const uint32_t spp = r->opts.samples;
    const uint32_t n_chunks = (spp + r->chunk_samples - 1) / r->chunk_samples;
    const bool reuse = n_chunks == 1 && r->primal_frame == (int64_t)frame;
    FrameParams fp = frame_params(r, frame);
    AdjointParams ap{d_image_dev, 1.0 / (double)spp, ws, n_materials, copies};
    uint64_t launches = 0;
    for (uint32_t c = 0; c < n_chunks; ++c) {
        const uint32_t s0 = c * r->chunk_samples;
        const uint32_t n = std::min(r->chunk_samples, spp - s0);
        if (!reuse) trace_chunk(r, fp, s0, n, nullptr, nullptr, &launches);
        trace_chunk(r, fp, s0, n, nullptr, nullptr, &launches, &ap);
    }
    launch_sum_grad(ws, 3 * n_materials, copies, grad_dev, r->stream);
    {
        const rayna_status ls = launch_status(r, "adjoint_static");
        if (ls != RAYNA_OK) return ls;
    }
    HIP_TRY_LAUNCHED(r, hipStreamSynchronize(r->stream));
    r->primal_frame = -1;
    return RAYNA_OK;
}

rayna_status rayna_primary_hits(rayna_renderer *r, uint32_t sample_index, int32_t *prim_rank, float *t) {
    if (!r || !prim_rank || !t) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    REQUIRE_LIVE(r);
    if (!r->cam_valid) return fail(r, RAYNA_INVALID_ARGUMENT, "camera is invalid");
    HIP_TRY(r, hipSetDevice(r->device));
    {
        const rayna_status ms = ensure_mt(r);
        if (ms != RAYNA_OK) return ms;
    }
    const uint32_t P = r->opts.width * r->opts.height;
    int32_t *rank_dev = nullptr;
    HIP_TRY(r, hipMalloc((void **)&rank_dev, (size_t)P * sizeof(int32_t)));
    FrameParams fp = frame_params(r, 0);
    fp.s0 = sample_index;
    fp.n_spp = 1;
    fp.n_slots = P;
    PathBuffers pb = r->pb;
    pb.hit_rank = rank_dev;
    r->primal_frame = -1;
    launch_generate(fp, pb, BatchView{}, r->stream);
    uint32_t *work = r->queue_count + 3 * (r->opts.ray_depth + 2);
    uint32_t *claims = r->claims + (size_t)(r->opts.ray_depth + 2) * EXTEND_CLAIM_WORDS;
    launch_fill_u32(work, 0u, 2, r->stream);
    launch_fill_u32(claims, 0u, EXTEND_CLAIM_WORDS, r->stream);
    launch_extend(r->dev.view, fp, pb, nullptr, nullptr, 0, r->grid_extend, r->dev.overflow, claims, r->tune,
                  BatchView{}, r->stream);
    launch_march(r->dev.view, fp, pb, nullptr, nullptr, (uint32_t)r->dev.view.n_sdfs, r->mt, r->mt_cap, work + 1,
                 r->grid_march, r->grid_shade, r->tune, BatchView{}, r->sdf_queue, r->sdf_count, r->stream);
    rayna_status ls = launch_status(r, "primary_hits");
    hipError_t e = hipSuccess;
    if (ls == RAYNA_OK) e = hipMemcpyAsync(prim_rank, rank_dev, (size_t)P * 4, hipMemcpyDeviceToHost, r->stream);
    if (ls == RAYNA_OK && e == hipSuccess)
        e = hipMemcpyAsync(t, pb.hit_t, (size_t)P * 4, hipMemcpyDeviceToHost, r->stream);
    // the stream has to drain before rank_dev is freed, whatever the launches reported
    const hipError_t se = hipStreamSynchronize(r->stream);
    if (e == hipSuccess) e = se;
    (void)hipFree(rank_dev);
    if (ls != RAYNA_OK) return ls;
    if (e != hipSuccess) {
        r->lost = std::string("primary hits: ") + hipGetErrorString(e);
        return fail(r, RAYNA_DEVICE_ERROR, r->lost);
    }
    return RAYNA_OK;
}

rayna_status rayna_dfs_order(rayna_renderer *r, int32_t *object, int32_t *local, int32_t capacity, int32_t *n_out) {
    if (!r || !n_out) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    const int32_t n = (int32_t)r->packed.rank_object.size();
    *n_out = n;
    if (object && local)
        for (int32_t i = 0; i < n && i < capacity; ++i) {
            object[i] = r->packed.rank_object[i];
            local[i] = r->packed.rank_local[i];
        }
    return RAYNA_OK;
}

rayna_status rayna_tree_hash(rayna_renderer *r, uint64_t *hash) {
    if (!r || !hash) return fail(r, RAYNA_INVALID_ARGUMENT, "null argument");
    *hash = r->packed.tree_hash;
    return RAYNA_OK;
}

rayna_status rayna_scene_prepare(const rayna_scene_desc *scene, uint64_t *tree_hash, int32_t *object,
                                 int32_t *local, int32_t capacity, int32_t *n_ranks, int32_t *stack_need) {
    g_create_error.clear();
    if (!scene) return fail(nullptr, RAYNA_INVALID_ARGUMENT, "null scene");
    PackedScene ps;
    std::string err;
    const rayna_status st = pack_scene(*scene, ps, err);
    if (st != RAYNA_OK) return fail(nullptr, st, err);
    if (tree_hash) *tree_hash = ps.tree_hash;
    if (n_ranks) *n_ranks = (int32_t)ps.rank_object.size();
    if (stack_need) *stack_need = ps.stack_need;
    if (object && local)
        for (int32_t i = 0; i < (int32_t)ps.rank_object.size() && i < capacity; ++i) {
            object[i] = ps.rank_object[i];
            local[i] = ps.rank_local[i];
        }
    return RAYNA_OK;
}

rayna_status rayna_scene_tree(const rayna_scene_desc *scene, int32_t *node_words, int32_t node_cap,
                              int32_t *item_words, int32_t item_cap, int32_t *xform_words, int32_t xform_cap,
                              int32_t counts[3], int32_t *root, int32_t *stack_need, int32_t *packet_need) {
    g_create_error.clear();
    if (!scene) return fail(nullptr, RAYNA_INVALID_ARGUMENT, "null scene");
    PackedScene ps;
    std::string err;
    const rayna_status st = pack_scene(*scene, ps, err);
    if (st != RAYNA_OK) return fail(nullptr, st, err);
    if (counts) {
        counts[0] = (int32_t)ps.nodes.size();
        counts[1] = (int32_t)ps.items.size();
        counts[2] = (int32_t)ps.xforms.size();
    }
    if (root) *root = ps.root;
    if (stack_need) *stack_need = ps.stack_need;
    if (packet_need) *packet_need = ps.packet_need;
    if (node_words)
        for (int32_t i = 0; i < (int32_t)ps.nodes.size() && i < node_cap; ++i)
            for (int k = 0; k < 4; ++k) {
                node_words[8 * i + k] = ps.nodes[i].child[k];
                node_words[8 * i + 4 + k] = ps.nodes[i].info[k];
            }
    if (item_words)
        for (int32_t i = 0; i < (int32_t)ps.items.size() && i < item_cap; ++i) {
            item_words[2 * i] = ps.items[i].kind;
            item_words[2 * i + 1] = ps.items[i].index;
        }
    if (xform_words)
        for (int32_t i = 0; i < (int32_t)ps.xforms.size() && i < xform_cap; ++i) {
            xform_words[2 * i] = ps.xforms[i].inner_kind;
            xform_words[2 * i + 1] = ps.xforms[i].inner_index;
        }
    return RAYNA_OK;
}

void rayna_destroy(rayna_renderer *r) {
    if (!r) return;
    (void)hipSetDevice(r->device);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    free_paths(r);
    free_scene(r->dev);
    if (r->grad) (void)hipFree(r->grad);
    if (r->grad_sum) (void)hipFree(r->grad_sum);
    if (r->mt) (void)hipFree(r->mt);
    if (r->sdf_buf) (void)hipFree(r->sdf_buf);
    if (r->rgba_dev) (void)hipFree(r->rgba_dev);
    if (r->checks) (void)hipFree(r->checks);
    for (hipEvent_t e : r->events) (void)hipEventDestroy(e);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

rayna_status rayna_debug_sqrt_mismatches(uint32_t first, uint64_t count, uint64_t *mismatches) {
    g_create_error.clear();
    if (!mismatches || count > (1ull << 32) - first)
        return fail(nullptr, RAYNA_INVALID_ARGUMENT, "null output or range beyond 2^32");
    *mismatches = 0;
    int device = 0;
    HIP_TRY(nullptr, hipGetDevice(&device));
    hipDeviceProp_t prop;
    HIP_TRY(nullptr, hipGetDeviceProperties(&prop, device));
    const uint32_t blocks = 16u * (uint32_t)std::max(1, prop.multiProcessorCount);
    unsigned long long *dev = nullptr;
    HIP_TRY(nullptr, hipMalloc((void **)&dev, sizeof(unsigned long long)));
    hipError_t e = hipMemset(dev, 0, sizeof(unsigned long long));
    // pieces of at most 2^31 patterns (the kernel's count is 32-bit)
    for (uint64_t done = 0; e == hipSuccess && done < count;) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(count - done, 1ull << 31);
        launch_sqrt_check((uint32_t)(first + done), n, dev, blocks, nullptr);
        e = hipGetLastError();
        done += n;
    }
    unsigned long long host = 0;
    if (e == hipSuccess) e = hipMemcpy(&host, dev, sizeof(host), hipMemcpyDeviceToHost);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(nullptr, RAYNA_DEVICE_ERROR, std::string("sqrt check: ") + hipGetErrorString(e));
    *mismatches = host;
    return RAYNA_OK;
}

rayna_status rayna_measure_valu_peak(uint32_t iters, double *wave_insts_per_s, double *seconds) {
    g_create_error.clear();
    if (!wave_insts_per_s || iters == 0 || iters > (1u << 20))
        return fail(nullptr, RAYNA_INVALID_ARGUMENT, "null output or iterations outside 1..2^20");
    int device = 0;
    HIP_TRY(nullptr, hipGetDevice(&device));
    hipDeviceProp_t prop;
    HIP_TRY(nullptr, hipGetDeviceProperties(&prop, device));
    // 32 blocks of four waves per CU: eight waves on each SIMD, four rounds of them
    const uint32_t blocks = 32u * (uint32_t)std::max(1, prop.multiProcessorCount);
    float *out = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipMalloc((void **)&out, (size_t)blocks * 256 * sizeof(float));
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    uint64_t insts = 0;
    float ms = 0.0f;
    if (e == hipSuccess) {
        launch_valu_peak(out, blocks, iters, nullptr);  // warm-up (clocks, code load)
        e = hipEventRecord(ev[0], nullptr);
    }
    if (e == hipSuccess) {
        insts = launch_valu_peak(out, blocks, iters, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return fail(nullptr, RAYNA_DEVICE_ERROR, std::string("VALU calibration: ") + hipGetErrorString(e));
    *wave_insts_per_s = ms > 0.0f ? (double)insts / (ms * 1e-3) : 0.0;
    if (seconds) *seconds = ms * 1e-3;
    return RAYNA_OK;
}

int32_t rayna_debug_fail_launches(int32_t countdown) { return g_fail_launches.exchange(countdown < 0 ? -1 : countdown); }

const char *rayna_last_error(const rayna_renderer *r) { return r ? r->error.c_str() : g_create_error.c_str(); }

}  // extern "C"
