// Batches of render requests (include/rayna_gpu.h: rayna_batch_*): several scenes merged into one device
// scene set, and the samples of many requests packed into chunks of paths that each run one wavefront
// pipeline (generate, extend and shade per depth, resolve) over all of their requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rayna_gpu.h"
#include "host_util.hpp"
#include "kernels.hpp"
#include "scene_build.hpp"

using namespace rayna;

namespace {

std::string g_batch_create_error;
// Paths in flight per chunk: at most 2^29 (four 1080p frames of 64 samples; ~50 GB of path state per pipeline,
// sized for the 288 GB of HBM of one MI355X).  Large chunks keep every depth's launch full and pay a depth's
// drain once for many requests.  A pipeline's buffers grow to the largest chunk of a call (a power of two).
constexpr uint64_t DEFAULT_CHUNK_SLOTS = 1ull << 29;
// Chunks in flight at once, each on its own stream with its own path state: while one chunk's depth drains
// (the last long paths of a launch hold a few waves), the other's launches fill the rest of the GPU.
constexpr int DEFAULT_PIPELINES = 3;
constexpr int MAX_PIPELINES = 4;

// Per-scene features that select kernel forms.
enum : uint32_t {
    SCENE_INSTANCES = 1u,
    SCENE_SPHERES = 2u,
    SCENE_MEDIA = 4u,
    SCENE_SDFS = 8u,
    SCENE_TEXTURED = 16u,
};

// One request's share of a chunk: samples [s0, s0 + n) of request `req`.
struct Piece {
    uint32_t req, s0, n;
};

// The chunks of one batch call and everything uploaded for them.
struct Plan {
    std::vector<std::vector<Piece>> chunks;
    std::vector<RequestParams> table;          // all chunks' entries, chunk c at [first_entry[c], ...)
    std::vector<uint32_t> first_entry;
    std::vector<uint32_t> finals;              // per chunk, entries (chunk-local) whose request ends there
    std::vector<uint32_t> first_final;         // chunk c's finals at [first_final[c], first_final[c + 1])
    std::vector<uint32_t> max_pixels;          // per chunk
    std::vector<uint32_t> resolve_blocks;      // per chunk: the most resolve blocks of one of its requests
    std::vector<uint32_t> depth;               // per chunk: deepest ray_depth of its requests
    std::vector<uint32_t> n_sdf;               // per chunk: most unbounded objects of its requests' scenes
    std::vector<uint32_t> features;            // per chunk: SCENE_* bits of any of its requests' scenes
    std::vector<uint32_t> slots;               // per chunk
    std::vector<int32_t> dep;                  // per chunk: the last earlier chunk holding a piece of one of its
                                               // requests (its resolve must come first), or -1
    std::vector<int32_t> last_chunk;           // per request (-1: not traced, e.g. invalid camera)
    std::vector<uint64_t> out_offset;          // per request, floats into the output buffer
    uint64_t out_floats = 0;
    uint32_t max_depth = 0;
    uint64_t key = 0;                          // hash of the request list (graph reuse)
};

// One chunk in flight: its stream and the path state of up to `capacity` paths.  Chunk c of a call runs on
// pipeline c % pipelines; a pipeline's chunks follow each other on its stream.
struct Pipeline {
    hipStream_t stream = nullptr;  // pipeline 0 runs on the batch's own stream
    hipEvent_t done = nullptr;     // after the pipeline's last chunk of a call (joins the batch's stream)
    uint64_t capacity = 0;
    std::vector<void *> allocs;
    PathBuffers pb{};
    uint32_t *slot_req = nullptr;
    int32_t *queue[2] = {nullptr, nullptr};
    uint32_t *queue_count = nullptr;
    uint32_t queue_count_len = 0;
    uint32_t *claims = nullptr;    // extend claim counters, EXTEND_CLAIM_WORDS per depth
    unsigned long long *segments = nullptr;
    float *mt = nullptr;           // march outcomes, 2 * capacity * n_sdf_max
    int32_t *sdf_queue = nullptr;  // scenes with unbounded objects: SDF-hit slots (normals in pb.sdf_n*)
    uint32_t *sdf_count = nullptr;
    int2 *overflow = nullptr;      // traversal stack entries beyond the LDS part
};

}  // namespace

struct rayna_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> scene_allocs;
    SceneView view{};
    std::vector<SceneSlot> slots;
    std::vector<uint32_t> scene_features;  // per scene: SCENE_* bits
    size_t overflow_entries = 0;  // per pipeline: extend grid * block * stack entries beyond the LDS part
    int cus = 1, grid_extend = 1024, grid_shade = 1024, grid_march = 1024;
    uint32_t n_sdf_max = 0;  // most unbounded objects of one scene
    uint64_t chunk_slots = DEFAULT_CHUNK_SLOTS;
    ExtendTuning tune;
    bool use_graph = false;
    int schedule = 0;
    int pipelines = DEFAULT_PIPELINES;
    Pipeline pipes[MAX_PIPELINES];
    hipEvent_t fork = nullptr;         // the call's start, which the other pipelines' streams wait for
    std::vector<hipEvent_t> resolved;  // per chunk: its resolve done (resolves run in chunk order)

    // per call
    RequestParams *table = nullptr;
    uint64_t table_cap = 0;
    uint32_t *finals = nullptr;
    uint64_t finals_cap = 0;
    float *sum = nullptr, *out = nullptr;
    uint64_t sum_cap = 0, out_cap = 0;
    std::vector<hipEvent_t> events;

    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
    std::string error;
};

namespace {

rayna_status bfail(rayna_batch *b, rayna_status st, const std::string &msg) {
    if (b) b->error = msg;
    else g_batch_create_error = msg;
    return st;
}

#define BHIP(b, expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return bfail((b), e_ == hipErrorOutOfMemory ? RAYNA_OUT_OF_MEMORY : RAYNA_DEVICE_ERROR,         \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                                 \
    } while (0)

template <class T>
rayna_status bupload(rayna_batch *b, const std::vector<T> &v, const T **out) {
    *out = nullptr;
    if (v.empty()) return RAYNA_OK;
    void *p = nullptr;
    BHIP(b, hipMalloc(&p, v.size() * sizeof(T)));
    b->scene_allocs.push_back(p);
    BHIP(b, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(p);
    return RAYNA_OK;
}

void free_graph(rayna_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_key = 0;
}

void free_pipe(rayna_batch *b, Pipeline &P) {
    for (void *p : P.allocs) (void)hipFree(p);
    P.allocs.clear();
    P.pb = PathBuffers{};
    P.slot_req = nullptr;
    P.queue[0] = P.queue[1] = nullptr;
    P.queue_count = nullptr;
    P.claims = nullptr;
    P.queue_count_len = 0;
    P.segments = nullptr;
    P.mt = nullptr;
    P.sdf_queue = nullptr;
    P.sdf_count = nullptr;
    P.overflow = nullptr;
    P.capacity = 0;
    free_graph(b);
}

void free_paths(rayna_batch *b) {
    for (Pipeline &P : b->pipes) free_pipe(b, P);
}

// Path state for `slots` paths (and its stream) in pipeline p.
rayna_status alloc_pipe(rayna_batch *b, int p, uint64_t slots, uint32_t max_depth) {
    Pipeline &P = b->pipes[p];
    if (!P.stream) {
        if (p == 0) P.stream = b->stream;
        else BHIP(b, hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    }
    if (!P.done) BHIP(b, hipEventCreateWithFlags(&P.done, hipEventDisableTiming));
    if (P.capacity >= slots && P.queue_count_len >= 3 * (max_depth + 2)) return RAYNA_OK;
    free_pipe(b, P);
    auto alloc = [&](size_t bytes, void **out) -> rayna_status {
        BHIP(b, hipMalloc(out, std::max<size_t>(bytes, 4)));
        P.allocs.push_back(*out);
        return RAYNA_OK;
    };
    rayna_status st;
    float **fields[] = {&P.pb.ox, &P.pb.oy, &P.pb.oz, &P.pb.dx, &P.pb.dy, &P.pb.dz, &P.pb.tr,
                        &P.pb.tg, &P.pb.tb, &P.pb.lr, &P.pb.lg, &P.pb.lb, &P.pb.hit_t};
    for (float **f : fields)
        if ((st = alloc(slots * sizeof(float), (void **)f)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&P.pb.hit_ref)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&P.pb.hit_aux)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&P.pb.hit_sub)) != RAYNA_OK) return st;
    if ((st = alloc(slots * sizeof(uint32_t), (void **)&P.slot_req)) != RAYNA_OK) return st;
    for (int q = 0; q < 2; ++q)
        if ((st = alloc(slots * sizeof(int32_t), (void **)&P.queue[q])) != RAYNA_OK) return st;
    // [0, D+2) paths entering each depth; [D+2, 2(D+2)) extend counters; [2(D+2), 3(D+2)) march counters
    const uint32_t qlen = 3 * (std::max<uint32_t>(max_depth, 16) + 2);
    if ((st = alloc(qlen * sizeof(uint32_t), (void **)&P.queue_count)) != RAYNA_OK) return st;
    if ((st = alloc((size_t)(qlen / 3) * EXTEND_CLAIM_WORDS * sizeof(uint32_t), (void **)&P.claims)) != RAYNA_OK)
        return st;
    if ((st = alloc(sizeof(unsigned long long), (void **)&P.segments)) != RAYNA_OK) return st;
    if (b->n_sdf_max) {
        if ((st = alloc(2 * slots * b->n_sdf_max * sizeof(float), (void **)&P.mt)) != RAYNA_OK) return st;
        float **planes[] = {&P.pb.sdf_nx, &P.pb.sdf_ny, &P.pb.sdf_nz};
        for (float **f : planes)
            if ((st = alloc(slots * sizeof(float), (void **)f)) != RAYNA_OK) return st;
        if ((st = alloc(slots * sizeof(int32_t), (void **)&P.sdf_queue)) != RAYNA_OK) return st;
        if ((st = alloc(sizeof(uint32_t), (void **)&P.sdf_count)) != RAYNA_OK) return st;
    }
    if (b->overflow_entries && (st = alloc(b->overflow_entries * sizeof(int2), (void **)&P.overflow)) != RAYNA_OK)
        return st;
    P.pb.hit_rank = nullptr;
    P.capacity = slots;
    P.queue_count_len = qlen;
    return RAYNA_OK;
}

template <class T>
rayna_status ensure(rayna_batch *b, T **p, uint64_t *cap, uint64_t n) {
    if (*cap >= n && *p) return RAYNA_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const uint64_t want = std::max<uint64_t>(n, 1);
    BHIP(b, hipMalloc((void **)p, want * sizeof(T)));
    *cap = want;
    free_graph(b);  // a captured graph holds the old pointers
    return RAYNA_OK;
}

// March items (slot, unbounded object) are counted in 32 bits.
void clamp_chunk(rayna_batch *b) {
    if (b->n_sdf_max && b->chunk_slots * b->n_sdf_max > 0xffffffffull) b->chunk_slots = 0xffffffffull / b->n_sdf_max;
}

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h;
}

rayna_status validate(rayna_batch *b, const rayna_request &q, int i) {
    const rayna_render_opts &o = q.opts;
    const std::string at = "request " + std::to_string(i) + ": ";
    if (q.scene < 0 || q.scene >= (int32_t)b->slots.size()) return bfail(b, RAYNA_INVALID_ARGUMENT, at + "bad scene");
    if (q.layout != RAYNA_LAYOUT_XMAJOR && q.layout != RAYNA_LAYOUT_ROWMAJOR)
        return bfail(b, RAYNA_INVALID_ARGUMENT, at + "unknown layout");
    if (o.width == 0 || o.height == 0 || o.samples == 0 || o.ray_branching == 0)
        return bfail(b, RAYNA_INVALID_ARGUMENT, at + "zero width, height, samples or branching");
    if (o.ray_branching != 1) return bfail(b, RAYNA_UNSUPPORTED, at + "ray_branching > 1");
    if (o.mode != RAYNA_MODE_PBR) return bfail(b, RAYNA_UNSUPPORTED, at + "only RenderMode::PBR");
    if (o.ray_depth > 1000) return bfail(b, RAYNA_UNSUPPORTED, at + "ray_depth above 1000");
    if ((uint64_t)o.width * o.height > b->chunk_slots)
        return bfail(b, RAYNA_UNSUPPORTED, at + "more pixels than paths per chunk");
    return RAYNA_OK;
}

// Cuts the requests' samples into chunks of at most chunk_slots paths, in request order and sample order, in
// whole sample layers (width * height paths each).
void make_plan(rayna_batch *b, const rayna_request *reqs, int n, const std::vector<GViewport> &vps,
               const std::vector<char> &valid, Plan &pl) {
    const uint64_t C = b->chunk_slots;
    pl = Plan{};
    pl.last_chunk.assign(n, -1);
    pl.out_offset.assign(n, 0);
    uint64_t fill = C;  // forces a new chunk
    int32_t prev_scene = -1;
    uint64_t h = 1469598103934665603ull;
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    const auto work = [&](int k) { return (uint64_t)reqs[k].opts.width * reqs[k].opts.height * reqs[k].opts.samples; };
    if (b->schedule == 1) {
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) < work(y); });
    } else if (b->schedule == 2 || b->schedule == 3) {
        std::vector<uint64_t> scene_work(b->slots.size(), 0);
        for (int i = 0; i < n; ++i) scene_work[reqs[i].scene] += work(i);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const int sx = reqs[x].scene, sy = reqs[y].scene;
            if (sx != sy) return scene_work[sx] != scene_work[sy] ? scene_work[sx] < scene_work[sy] : sx < sy;
            return work(x) < work(y);
        });
    }
    h = mix(h, (uint64_t)b->schedule);
    for (int oi = 0; oi < n; ++oi) {
        const int i = order[oi];
        const rayna_request &q = reqs[i];
        const uint64_t P = (uint64_t)q.opts.width * q.opts.height;
        pl.out_offset[i] = pl.out_floats;
        pl.out_floats += 3 * P;
        h = mix(h, ((uint64_t)q.scene << 48) ^ ((uint64_t)q.opts.width << 32) ^ ((uint64_t)q.opts.height << 16) ^
                       q.opts.samples);
        h = mix(h, q.opts.seed);
        h = mix(h, ((uint64_t)q.opts.ray_depth << 8) ^ (uint64_t)q.layout);
        uint64_t cam_bits[9];
        std::memcpy(cam_bits, &q.camera, sizeof(cam_bits));
        for (uint64_t w : cam_bits) h = mix(h, w);
        if (!valid[i]) continue;
        if (b->schedule == 3 && q.scene != prev_scene) fill = C;  // interleaved: chunks hold one scene each
        prev_scene = q.scene;
        uint32_t done = 0;
        while (done < q.opts.samples) {
            uint64_t fit = (C - fill) / P;
            if (fit == 0) {
                pl.chunks.emplace_back();
                fill = 0;
                fit = C / P;
            }
            const uint32_t take = (uint32_t)std::min<uint64_t>(fit, q.opts.samples - done);
            pl.chunks.back().push_back(Piece{(uint32_t)i, done, take});
            fill += take * P;
            done += take;
        }
        pl.last_chunk[i] = (int32_t)pl.chunks.size() - 1;
    }
    pl.key = h;
    if (b->schedule == 3 && !pl.chunks.empty()) {
        // interleaved: the by-scene chunks dealt round-robin over the scene groups (groups in the by-scene
        // order), so that the chunks in flight on the pipelines at one time mostly belong to different scenes
        std::vector<std::vector<std::vector<Piece>>> groups;
        int32_t last_scene = -1;
        for (auto &ch : pl.chunks) {
            const int32_t sc = reqs[ch.front().req].scene;
            if (groups.empty() || sc != last_scene) groups.emplace_back();
            groups.back().push_back(std::move(ch));
            last_scene = sc;
        }
        pl.chunks.clear();
        for (size_t k = 0;; ++k) {
            bool any = false;
            for (auto &g : groups)
                if (k < g.size()) {
                    pl.chunks.push_back(std::move(g[k]));
                    any = true;
                }
            if (!any) break;
        }
        for (size_t c = 0; c < pl.chunks.size(); ++c)
            for (const Piece &pc : pl.chunks[c])
                if (pc.s0 + pc.n == reqs[pc.req].opts.samples) pl.last_chunk[pc.req] = (int32_t)c;
    }
    {
        std::vector<int32_t> seen(n, -1);  // per request: the last chunk so far with a piece of it
        for (size_t c = 0; c < pl.chunks.size(); ++c) {
            int32_t d = -1;
            for (const Piece &pc : pl.chunks[c]) {
                d = std::max(d, seen[pc.req]);
                seen[pc.req] = (int32_t)c;
            }
            pl.dep.push_back(d);
        }
    }
    for (size_t c = 0; c < pl.chunks.size(); ++c) {
        pl.first_entry.push_back((uint32_t)pl.table.size());
        pl.first_final.push_back((uint32_t)pl.finals.size());
        uint32_t slot = 0, maxp = 0, depth = 0, nsdf = 0, rblocks = 1, feat = 0;
        for (size_t e = 0; e < pl.chunks[c].size(); ++e) {
            const Piece &pc = pl.chunks[c][e];
            const rayna_request &q = reqs[pc.req];
            const uint32_t P = q.opts.width * q.opts.height;
            RequestParams R{};
            R.vp = vps[pc.req];
            R.width = q.opts.width;
            R.height = q.opts.height;
            R.spp = q.opts.samples;
            R.sdim = isqrt(q.opts.samples);
            R.key0 = (uint32_t)q.opts.seed;
            R.key1 = (uint32_t)(q.opts.seed >> 32);  // frame 0
            R.ray_depth = q.opts.ray_depth;
            R.n_pixels = P;
            R.s0 = pc.s0;
            R.n_spp = pc.n;
            R.slot_base = slot;
            R.n_slots = pc.n * P;
            const SceneSlot &sc = b->slots[q.scene];
            R.root = sc.root;
            R.first_sdf = sc.first_sdf;
            R.n_sdfs = sc.n_sdfs;
            R.skybox = sc.skybox;
            R.sky_image = sc.sky_image;
            R.sum_offset = (uint32_t)pl.out_offset[pc.req];
            R.layout = q.layout;
            R.first_piece = pc.s0 == 0 ? 1u : 0u;
            pl.table.push_back(R);
            if (pc.s0 + pc.n == q.opts.samples) pl.finals.push_back((uint32_t)e);
            slot += R.n_slots;
            maxp = std::max(maxp, P);
            const uint32_t per = resolve_pixels_per_block(pc.n);
            rblocks = std::max(rblocks, (P + per - 1) / per);
            depth = std::max(depth, q.opts.ray_depth);
            nsdf = std::max(nsdf, (uint32_t)sc.n_sdfs);
            feat |= b->scene_features[q.scene];
        }
        pl.max_pixels.push_back(maxp);
        pl.resolve_blocks.push_back(rblocks);
        pl.depth.push_back(depth);
        pl.n_sdf.push_back(nsdf);
        pl.features.push_back(feat);
        pl.slots.push_back(slot);
        pl.max_depth = std::max(pl.max_depth, depth);
    }
    pl.first_final.push_back((uint32_t)pl.finals.size());
}

// The merged scene set's view as one chunk sees it: the kernels compile out the sphere test, media, instance
// and texture code and, without SDFs, the SDF normal code (k_shade's eight-wave form), so a chunk whose requests'
// scenes hold none of these runs the forms a single render of those scenes would.  A request's rays only reach
// its own scene's tree, so the other scenes' items never meet the compiled-out code.
SceneView chunk_view(const SceneView &merged, uint32_t features) {
    SceneView v = merged;
    v.instances = (features & SCENE_INSTANCES) ? merged.instances : 0;
    v.sphere_tests = (features & SCENE_SPHERES) ? merged.sphere_tests : 0;
    v.media = (features & SCENE_MEDIA) ? merged.media : 0;
    v.textured = (features & SCENE_TEXTURED) ? merged.textured : 0;
    v.n_sdfs = (features & SCENE_SDFS) ? merged.n_sdfs : 0;
    return v;
}

// Launches of a plan traced on `used` pipelines.
uint32_t plan_launches(const Plan &pl, int used) {
    uint32_t n = (uint32_t)used;  // segment counters
    for (size_t c = 0; c < pl.chunks.size(); ++c) {
        const uint32_t nf = pl.first_final[c + 1] - pl.first_final[c];
        n += 3 + (pl.n_sdf[c] ? 6 : 2) * (pl.depth[c] + 1) + 2 + (nf ? 1 : 0);
    }
    return n;
}

// Enqueues every chunk of the plan, chunk c on pipeline c % used.  The pipelines' streams start after the
// batch's stream (which holds the uploads) and the batch's stream waits for all of them at the end.  A chunk's
// resolve waits for the resolve of the last earlier chunk holding a piece of one of its requests, so every
// request's sums are added in sample order, as on one pipeline.
// `timing` (eager runs only): events[0] at the start and events[1 + c] after chunk c.
hipError_t enqueue(rayna_batch *b, const Plan &pl, int used, bool timing) {
    hipError_t err = hipSuccess;
    auto check = [&](hipError_t e) {
        if (err == hipSuccess) err = e;
    };
    hipStream_t s0 = b->stream;
    if (timing) check(hipEventRecord(b->events[0], s0));
    if (used > 1) {
        check(hipEventRecord(b->fork, s0));
        for (int p = 1; p < used; ++p) check(hipStreamWaitEvent(b->pipes[p].stream, b->fork, 0));
    }
    for (int p = 0; p < used; ++p)
        launch_fill_u32(reinterpret_cast<uint32_t *>(b->pipes[p].segments), 0u, 2, b->pipes[p].stream);
    for (size_t c = 0; c < pl.chunks.size(); ++c) {
        Pipeline &P = b->pipes[c % used];
        hipStream_t s = P.stream;
        BatchView bv{b->table + pl.first_entry[c], (uint32_t)pl.chunks[c].size(), P.slot_req};
        FrameParams fp{};
        fp.n_slots = pl.slots[c];
        fp.ray_depth = pl.depth[c];
        const uint32_t D = pl.depth[c];
        const SceneView cv = chunk_view(b->view, pl.features[c]);
        launch_fill_u32(P.queue_count, 0u, 3 * (D + 2), s);
        launch_fill_u32(P.claims, 0u, (D + 1) * EXTEND_CLAIM_WORDS, s);
        // the camera rays are stored only where a kernel loads them: k_extend and k_march at depth 0 (the camera-ray
        // packets and the first k_shade compute them); k_generate records every slot's request either way
        fp.rays_stored = (pl.features[c] & SCENE_SDFS) != 0u || !extend_packets(cv, fp, b->tune, bv, 0);
        launch_generate(fp, P.pb, bv, s);
        for (uint32_t depth = 0; depth <= D; ++depth) {
            const int32_t *qin = depth == 0 ? nullptr : P.queue[depth & 1];
            const uint32_t *cin = depth == 0 ? nullptr : P.queue_count + depth;
            int32_t *qout = P.queue[(depth + 1) & 1];
            uint32_t *cout = P.queue_count + depth + 1;
            uint32_t *work = P.claims + (size_t)depth * EXTEND_CLAIM_WORDS;
            launch_extend(cv, fp, P.pb, qin, cin, depth, b->grid_extend, P.overflow, work, b->tune, bv, s);
            launch_march(cv, fp, P.pb, qin, cin, pl.n_sdf[c], P.mt, P.capacity * b->n_sdf_max,
                         P.queue_count + 2 * (D + 2) + depth, b->grid_march, b->grid_shade, b->tune, bv,
                         P.sdf_queue, P.sdf_count, s);
            launch_shade(cv, fp, P.pb, qin, cin, qout, cout, depth, b->grid_shade, nullptr, bv, s);
        }
        launch_count_segments(P.queue_count, D + 1, fp.n_slots, P.segments, s);
        // (a dependency on the same pipeline is kept by its stream's order)
        if (used > 1 && pl.dep[c] >= 0 && (size_t)pl.dep[c] % used != c % used)
            check(hipStreamWaitEvent(s, b->resolved[pl.dep[c]], 0));
        launch_resolve_batch(bv, pl.resolve_blocks[c], P.pb, b->sum, s);
        const uint32_t nf = pl.first_final[c + 1] - pl.first_final[c];
        launch_finalize_batch(bv, b->finals + pl.first_final[c], nf, pl.max_pixels[c], b->sum, b->out, s);
        if (used > 1) check(hipEventRecord(b->resolved[c], s));
        if (timing) check(hipEventRecord(b->events[1 + c], s));
    }
    for (int p = 1; p < used; ++p) {
        check(hipEventRecord(b->pipes[p].done, b->pipes[p].stream));
        check(hipStreamWaitEvent(s0, b->pipes[p].done, 0));
    }
    return err;
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

rayna_status rayna_batch_create(const rayna_scene_desc *scenes, int32_t n_scenes, rayna_batch **out) {
    g_batch_create_error.clear();
    if (!scenes || n_scenes <= 0 || !out) return bfail(nullptr, RAYNA_INVALID_ARGUMENT, "no scenes or null out");
    *out = nullptr;
    rayna_batch *b = new rayna_batch();
    auto bail = [&](rayna_status st) {
        g_batch_create_error = b->error;
        rayna_batch_destroy(b);
        return st;
    };
    if (hipGetDevice(&b->device) != hipSuccess) {
        b->error = "no HIP device";
        return bail(RAYNA_DEVICE_ERROR);
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->error = "hipStreamCreate failed";
        return bail(RAYNA_DEVICE_ERROR);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) {
        b->error = "hipGetDeviceProperties failed";
        return bail(RAYNA_DEVICE_ERROR);
    }
    b->cus = std::max(1, prop.multiProcessorCount);
    b->grid_extend = b->cus * extend_blocks_per_cu();
    b->grid_march = b->cus * march_blocks_per_cu();
    b->grid_shade = b->cus * SHADE_BLOCKS_PER_CU;
    std::vector<PackedScene> packed(n_scenes);
    b->scene_features.assign(n_scenes, 0u);
    for (int i = 0; i < n_scenes; ++i) {
        std::string err;
        const rayna_status st = pack_scene(scenes[i], packed[i], err);
        if (st != RAYNA_OK) {
            b->error = "scene " + std::to_string(i) + ": " + err;
            return bail(st);
        }
    }
    for (int i = 0; i < n_scenes; ++i) {
        const PackedScene &p = packed[i];
        uint32_t f = 0;
        if (has_instances(p)) f |= SCENE_INSTANCES;
        if (has_spheres(p)) f |= SCENE_SPHERES;
        if (!p.volumes.empty()) f |= SCENE_MEDIA;
        if (!p.sdfs.empty()) f |= SCENE_SDFS;
        if (has_textures(p) || p.skybox == RAYNA_SKY_HDRI) f |= SCENE_TEXTURED;
        b->scene_features[i] = f;
    }
    PackedScene merged;
    if (!merge_scenes(packed, merged, b->slots)) {
        b->error = "the scenes together exceed the index range of the device layout";
        return bail(RAYNA_UNSUPPORTED);
    }
    if (merged.stack_need > 4096) {
        b->error = "scene BVH is deeper than 4096 stack entries";
        return bail(RAYNA_UNSUPPORTED);
    }
    rayna_status st;
    SceneView &v = b->view;
#define UP(vec, field) \
    if ((st = bupload(b, merged.vec, &v.field)) != RAYNA_OK) return bail(st);
    merged.nodes = device_nodes(merged.nodes);
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
    for (const SceneSlot &sl : b->slots) b->n_sdf_max = std::max<uint32_t>(b->n_sdf_max, (uint32_t)sl.n_sdfs);
    clamp_chunk(b);
    v.root = -1;  // per request
    v.n_sdfs = (int32_t)merged.sdfs.size();
    v.skybox = RAYNA_SKY_NONE;
    v.stack_need = merged.stack_need;
    v.packet_need = merged.packet_need;
    v.instances = has_instances(merged) ? 1 : 0;
    v.sphere_tests = has_spheres(merged) ? 1 : 0;
    v.media = merged.volumes.empty() ? 0 : 1;
    v.textured = has_textures(merged) ? 1 : 0;
    for (const SceneSlot &sl : b->slots)
        if (sl.skybox == RAYNA_SKY_HDRI) v.textured = 1;
    const size_t deep = (size_t)std::max(0, merged.stack_need - extend_lds_stack());
    b->overflow_entries = deep * (size_t)b->grid_extend * extend_block();
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        b->error = "hipEventCreate failed";
        return bail(RAYNA_DEVICE_ERROR);
    }
    *out = b;
    return RAYNA_OK;
}

rayna_status rayna_batch_set_tuning(rayna_batch *b, const rayna_tuning *t) {
    if (!b || !t) return bfail(b, RAYNA_INVALID_ARGUMENT, "null argument");
    if (t->refill_lanes > 64 || t->wave_batch > 4096 || t->shade_blocks_per_cu > 64 ||
        (t->chunk_slots && t->chunk_slots < (1u << 16)) || t->schedule < 0 || t->schedule > 3 ||
        t->pipelines < 0 || t->pipelines > MAX_PIPELINES)
        return bfail(b, RAYNA_INVALID_ARGUMENT, "tuning out of range");
    if (t->use_graph && t->schedule == 3)
        return bfail(b, RAYNA_UNSUPPORTED, "the interleaved schedule is not captured in a graph");
    b->tune.refill_lanes = t->refill_lanes ? t->refill_lanes : ExtendTuning{}.refill_lanes;
    b->tune.wave_batch = t->wave_batch ? t->wave_batch : ExtendTuning{}.wave_batch;
    if (t->packet_depths < -1 || t->packet_depths > 16) return bfail(b, RAYNA_INVALID_ARGUMENT, "tuning out of range");
    b->tune.packet_depths = t->packet_depths == 0 ? ExtendTuning{}.packet_depths
                                                   : (uint32_t)std::max(0, t->packet_depths);
    b->grid_shade = b->cus * (int)(t->shade_blocks_per_cu ? t->shade_blocks_per_cu : SHADE_BLOCKS_PER_CU);
    const uint64_t chunk = b->chunk_slots;
    b->chunk_slots = t->chunk_slots ? t->chunk_slots : DEFAULT_CHUNK_SLOTS;
    clamp_chunk(b);
    if (b->chunk_slots != chunk) free_paths(b);
    b->pipelines = t->pipelines ? t->pipelines : DEFAULT_PIPELINES;
    for (int p = b->pipelines; p < MAX_PIPELINES; ++p) free_pipe(b, b->pipes[p]);
    b->use_graph = t->use_graph != 0;
    b->schedule = t->schedule;
    free_graph(b);
    return RAYNA_OK;
}

rayna_status rayna_render_batch(rayna_batch *b, const rayna_request *reqs, int32_t n, rayna_batch_stats *stats) {
    if (!b) return bfail(b, RAYNA_INVALID_ARGUMENT, "null batch");
    if (n < 0 || (n > 0 && !reqs)) return bfail(b, RAYNA_INVALID_ARGUMENT, "bad request list");
    if (n > 65535) return bfail(b, RAYNA_UNSUPPORTED, "more than 65535 requests in one batch");
    const auto t0 = std::chrono::steady_clock::now();
    BHIP(b, hipSetDevice(b->device));
    std::vector<GViewport> vps(n);
    std::vector<char> valid(n, 0);
    for (int i = 0; i < n; ++i) {
        const rayna_status st = validate(b, reqs[i], i);
        if (st != RAYNA_OK) return st;
        valid[i] = viewport(reqs[i].camera, vps[i]) ? 1 : 0;
    }
    Plan pl;
    make_plan(b, reqs, n, vps, valid, pl);
    rayna_status st;
    const int used = (int)std::max<size_t>(1, std::min<size_t>(b->pipelines, pl.chunks.size()));
    uint64_t need = 1ull << 16;
    for (uint32_t sl : pl.slots)
        while (need < sl) need <<= 1;
    need = std::min(need, b->chunk_slots);
    for (int p = 0; p < used; ++p)
        if ((st = alloc_pipe(b, p, need, pl.max_depth)) != RAYNA_OK) return st;
    if ((st = ensure(b, &b->table, &b->table_cap, pl.table.size())) != RAYNA_OK) return st;
    if ((st = ensure(b, &b->finals, &b->finals_cap, pl.finals.size())) != RAYNA_OK) return st;
    if (pl.out_floats >= (1ull << 32))
        return bfail(b, RAYNA_UNSUPPORTED, "the batch's images exceed 2^32 floats; split it");
    if ((st = ensure(b, &b->sum, &b->sum_cap, pl.out_floats)) != RAYNA_OK) return st;
    if ((st = ensure(b, &b->out, &b->out_cap, pl.out_floats)) != RAYNA_OK) return st;
    if (!pl.table.empty())
        BHIP(b, hipMemcpyAsync(b->table, pl.table.data(), pl.table.size() * sizeof(RequestParams),
                               hipMemcpyHostToDevice, b->stream));
    if (!pl.finals.empty())
        BHIP(b, hipMemcpyAsync(b->finals, pl.finals.data(), pl.finals.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, b->stream));
    const bool graph = b->use_graph;
    const size_t n_events = pl.chunks.size() + 2;
    while (b->events.size() < n_events) {
        hipEvent_t e;
        BHIP(b, hipEventCreate(&e));
        b->events.push_back(e);
    }
    while (b->resolved.size() < pl.chunks.size()) {
        hipEvent_t e;
        BHIP(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->resolved.push_back(e);
    }
    const uint32_t launches = plan_launches(pl, used);
    if (graph) {
        const uint64_t key = mix(pl.key, (uint64_t)used);
        if (!b->graph_exec || b->graph_key != key) {
            free_graph(b);
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t eq = enqueue(b, pl, used, false);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (eq != hipSuccess || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, RAYNA_DEVICE_ERROR,
                             std::string("graph capture: ") + hipGetErrorString(eq != hipSuccess ? eq : e));
            }
            b->graph = g;
            BHIP(b, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            b->graph_key = key;
        }
        BHIP(b, hipEventRecord(b->events[0], b->stream));
        BHIP(b, hipGraphLaunch(b->graph_exec, b->stream));
        BHIP(b, hipEventRecord(b->events[1], b->stream));
    } else {
        BHIP(b, enqueue(b, pl, used, true));
    }
    BHIP(b, hipGetLastError());
    for (int i = 0; i < n; ++i) {
        if (!reqs[i].out_rgb || !valid[i]) continue;
        BHIP(b, hipMemcpyAsync(reqs[i].out_rgb, b->out + pl.out_offset[i],
                               3 * (size_t)reqs[i].opts.width * reqs[i].opts.height * sizeof(float),
                               hipMemcpyDeviceToHost, b->stream));
    }
    unsigned long long seg[MAX_PIPELINES] = {};
    for (int p = 0; p < used; ++p)
        BHIP(b, hipMemcpyAsync(&seg[p], b->pipes[p].segments, sizeof(seg[p]), hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    unsigned long long segs = 0;
    for (int p = 0; p < used; ++p) segs += seg[p];
    for (int i = 0; i < n; ++i) {
        if (valid[i] || !reqs[i].out_rgb) continue;
        write_checkerboard(ilog2(reqs[i].opts.width), ilog2(reqs[i].opts.height), reqs[i].out_rgb, reqs[i].layout);
    }
    if (stats) {
        rayna_batch_stats s{};
        s.request_ms = stats->request_ms;
        float ms = 0;
        if (graph) {
            BHIP(b, hipEventElapsedTime(&ms, b->events[0], b->events[1]));
        } else {  // the chunks may end out of order on the pipelines
            for (size_t c = 0; c < pl.chunks.size(); ++c) {
                float m = 0;
                BHIP(b, hipEventElapsedTime(&m, b->events[0], b->events[1 + c]));
                ms = std::max(ms, m);
            }
        }
        s.kernel_ms = ms;
        for (int i = 0; i < n; ++i) {
            if (valid[i]) s.samples += (uint64_t)reqs[i].opts.width * reqs[i].opts.height * reqs[i].opts.samples;
            if (!s.request_ms) continue;
            s.request_ms[i] = 0.0;
            if (!graph && pl.last_chunk[i] >= 0) {
                float m = 0;
                BHIP(b, hipEventElapsedTime(&m, b->events[0], b->events[1 + pl.last_chunk[i]]));
                s.request_ms[i] = m;
            }
        }
        s.segments = segs;
        s.chunks = (uint32_t)pl.chunks.size();
        s.launches = launches;
        s.duration_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *stats = s;
    }
    return RAYNA_OK;
}

void rayna_batch_destroy(rayna_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Pipeline &P : b->pipes)
        if (P.stream && P.stream != b->stream) (void)hipStreamSynchronize(P.stream);
    free_paths(b);
    for (Pipeline &P : b->pipes) {
        if (P.stream && P.stream != b->stream) (void)hipStreamDestroy(P.stream);
        if (P.done) (void)hipEventDestroy(P.done);
    }
    for (hipEvent_t e : b->resolved) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (void *p : b->scene_allocs) (void)hipFree(p);
    if (b->table) (void)hipFree(b->table);
    if (b->finals) (void)hipFree(b->finals);
    if (b->sum) (void)hipFree(b->sum);
    if (b->out) (void)hipFree(b->out);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *rayna_batch_last_error(const rayna_batch *b) { return b ? b->error.c_str() : g_batch_create_error.c_str(); }

}  // extern "C"
