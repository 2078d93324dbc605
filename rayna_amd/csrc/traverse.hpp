// Closest-hit traversal of the extend kernel (included by kernels.hip inside its anonymous namespace, after the
// primitive tests).
//
// Result: the closest hit over the BVH with ties to the higher depth-first rank -- what the engine's
// shrinking inclusive interval computes (mesh/advanced/bvh.rs:77-91, object/bvh.rs:70-108).  The unbounded
// objects follow in march.hpp.  Culling is conservative, so the result does not depend on the order nodes are
// visited in.
//
// Organisation (Aila & Laine, "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009):
//  * while-while: leaf ranges are pushed on the stack like inner nodes; a lane first pops inner nodes until
//    it reaches a leaf, then tests that leaf's primitives one per iteration, so the lanes of a wave run the
//    box code together and the primitive code together;
//  * speculative inner loop: a lane that has found a leaf range keeps expanding inner nodes until every lane of
//    its wave holds one, so the primitive loop runs with more lanes busy;
//  * persistent waves with dynamic fetch: each wave takes rays from the launch's claim counters whenever enough
//    of its lanes are idle, so lanes whose rays finish early pick up new rays instead of waiting.
//
// Stack entry codes and the node words that hold them are described in layout.hpp.

// Traversal counters, compiled in only for measurement builds (-DRAYNA_TRAVERSAL_STATS): per lane, node pops,
// node expansions, primitive tests and loop iterations, summed into a device array read by
// rayna_traversal_stats().
#ifdef RAYNA_TRAVERSAL_STATS
__device__ unsigned long long g_trav_stats[4];
#define RAYNA_STAT(name, v) (stat_##name += (v))
#define RAYNA_STAT_DECL() unsigned long long stat_pops = 0, stat_expands = 0, stat_prims = 0, stat_iterations = 0
#define RAYNA_STAT_FLUSH()                            \
    do {                                              \
        atomicAdd(&g_trav_stats[0], stat_pops);       \
        atomicAdd(&g_trav_stats[1], stat_expands);    \
        atomicAdd(&g_trav_stats[2], stat_prims);      \
        atomicAdd(&g_trav_stats[3], stat_iterations); \
    } while (0)
#else
#define RAYNA_STAT(name, v) ((void)0)
#define RAYNA_STAT_DECL() ((void)0)
#define RAYNA_STAT_FLUSH() ((void)0)
#endif


// Traversal stack of one lane: the top EXT_LDS_STACK entries in LDS, each (code, distance) one 64-bit word (one
// column per lane, conflict-free), deeper entries in this thread's column of a global overflow array (`stride` =
// threads in the grid).  The host sizes the overflow for the deepest stack the tree can need, so no push is ever
// dropped.  The LDS column is addressed through an LDS pointer and each entry moves as one 64-bit word: through
// generic pointers the compiler merged the LDS and global halves of a push into one flat store.
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

__device__ __forceinline__ unsigned long long stack_word(int32_t n, float t) {
    return (unsigned long long)(uint32_t)n | ((unsigned long long)__float_as_uint(t) << 32);
}

struct LaneStack {
    lds_u64 *s_entry;
    int2 *g;
    uint32_t stride;
    int sp;
    // The top entry stays in registers until the next push: the nearest child an expansion pushes last is popped
    // next without touching LDS.
    int32_t top_n = 0;
    float top_t = 0.0f;
    bool top_valid = false;
    __device__ __forceinline__ void store(int i, int32_t n, float t) {
        if (i < EXT_LDS_STACK) s_entry[i * EXT_BLOCK + threadIdx.x] = stack_word(n, t);
        else g[(size_t)(i - EXT_LDS_STACK) * stride] = make_int2(n, __float_as_int(t));
    }
    __device__ __forceinline__ void push(int32_t n, float t) {
        if (top_valid) store(sp - 1, top_n, top_t);
        top_n = n;
        top_t = t;
        top_valid = true;
        ++sp;
    }
    __device__ __forceinline__ void pop(int32_t *n, float *t) {
        --sp;
        if (top_valid) {
            top_valid = false;
            *n = top_n;
            *t = top_t;
            return;
        }
        if (sp < EXT_LDS_STACK) {
            const unsigned long long e = s_entry[sp * EXT_BLOCK + threadIdx.x];
            *n = (int32_t)(uint32_t)e;
            *t = __uint_as_float((uint32_t)(e >> 32));
        } else {
            const int2 e = g[(size_t)(sp - EXT_LDS_STACK) * stride];
            *n = e.x;
            *t = __int_as_float(e.y);
        }
    }
};

// Tests the four child boxes of inner node `node` against [T_MIN, best] (conservatively) and pushes the hit
// children -- inner nodes and leaf ranges alike -- far to near, so the nearest is popped first.
__device__ __forceinline__ void expand(const SceneView &sv, const DRay &ray, float best, int32_t node, LaneStack &st) {
    const float4 *np = reinterpret_cast<const float4 *>(sv.nodes + node);
    const float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5];
    const int4 cc = reinterpret_cast<const int4 *>(np)[7];  // the slots' entry codes (device node copy)
    const float lo_x[4] = {lx.x, lx.y, lx.z, lx.w}, lo_y[4] = {ly.x, ly.y, ly.z, ly.w},
                lo_z[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hi_x[4] = {hx.x, hx.y, hx.z, hx.w}, hi_y[4] = {hy.x, hy.y, hy.z, hy.w},
                hi_z[4] = {hz.x, hz.y, hz.z, hz.w};
    int32_t code[4] = {cc.x, cc.y, cc.z, cc.w};
    float key[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float tx1 = (lo_x[k] - ray.o.x) * ray.inv.x, tx2 = (hi_x[k] - ray.o.x) * ray.inv.x;
        float t0 = fminf(tx1, tx2), t1 = fmaxf(tx1, tx2);
        const float ty1 = (lo_y[k] - ray.o.y) * ray.inv.y, ty2 = (hi_y[k] - ray.o.y) * ray.inv.y;
        t0 = fmaxf(t0, fminf(ty1, ty2));
        t1 = fminf(t1, fmaxf(ty1, ty2));
        const float tz1 = (lo_z[k] - ray.o.z) * ray.inv.z, tz2 = (hi_z[k] - ray.o.z) * ray.inv.z;
        t0 = fmaxf(t0, fminf(tz1, tz2));
        t1 = fminf(t1, fmaxf(tz1, tz2));
        const float lo = fmaxf(t0, T_MIN), hi = fminf(t1, best);
        const bool hit = code[k] != ENTRY_NONE && lo <= hi * BOX_SLACK;
        key[k] = hit ? t0 : -__builtin_inff();
        if (!hit) code[k] = ENTRY_NONE;  // not pushed
    }
#define RAYNA_CSWAP(a, b)              \
    if (key[a] < key[b]) {             \
        const float tt_ = key[a];      \
        key[a] = key[b];               \
        key[b] = tt_;                  \
        const int32_t cc_ = code[a];   \
        code[a] = code[b];             \
        code[b] = cc_;                 \
    }
    RAYNA_CSWAP(0, 1)
    RAYNA_CSWAP(2, 3)
    RAYNA_CSWAP(0, 2)
    RAYNA_CSWAP(1, 3)
    RAYNA_CSWAP(1, 2)
#undef RAYNA_CSWAP
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (code[k] != ENTRY_NONE) st.push(code[k], key[k]);
}

// Instances (an object's mesh BVH under a non-identity transform, SimpleObject::full_intersect around
// ObjectTransform, object/transform.rs): the lane traverses the instanced tree in object space with a closest hit
// of its own, found as BvhMesh::intersect finds it (mesh/advanced/bvh.rs:63-102) from an interval [T_MIN, +inf).
// Entering pushes the rest of the current leaf range, the closest hit so far (three words), a marker -- the
// instance's item index with a NaN distance -- and the instanced root; the lane's ray becomes the object-space
// ray.  When the marker is popped every entry of the instance has been, and the instance's closest hit, mapped
// out (its distance measured from the world ray's origin), competes with the saved one.
__device__ __forceinline__ void enter_instance(const GXform *x, int32_t item, int32_t code, int i, float leaf_t,
                                               DRay &ray, Best &b, LaneStack &st) {
    const int32_t start = (int32_t)((uint32_t)code & LEAF_CODE_START);
    const int n = (int)(((uint32_t)code >> 27) & 7u) + 1;
    if (i + 1 < n) st.push(leaf_code(start + i + 1, n - i - 1, true), leaf_t);
    st.push((int32_t)b.ref, b.t);
    st.push(b.rank, __uint_as_float(b.aux));
    st.push((int32_t)b.sub, 0.0f);
    st.push(item, __builtin_nanf(""));
    st.push(x->inner_index, -__builtin_inff());
    ray = object_ray(x, ray);
    b = Best{FLT_MAX, -1, 0u, 0u, 0u};
}

__device__ __forceinline__ void finish_instance(const SceneView &sv, const PathBuffers &pb, uint32_t slot,
                                                int32_t item, DRay &ray, Best &b, LaneStack &st) {
    int32_t sub, rank, ref;
    float unused, aux, t;
    st.pop(&sub, &unused);
    st.pop(&rank, &aux);
    st.pop(&ref, &t);
    Best w{t, rank, (uint32_t)ref, __float_as_uint(aux), (uint32_t)sub};
    const DRay world = ray_from_dir(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]},
                                    f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
    if (b.ref != 0u) {
        const GXform *x = sv.xforms + sv.items[item].index;
        const float tw = length(affine_point(x->fwd, at(ray, b.t)) - world.o);
        if (tw <= w.t && (tw < w.t || b.rank > w.rank))
            w = Best{tw, b.rank, make_ref(ITEM_XFORM, item), __float_as_uint(b.t), b.ref};
    }
    b = w;
    ray = world;
}

// Tests entry `i` of leaf range `code` and keeps it if closer, or equally close with a higher rank.  Returns true
// when the entry is an instance, which ends the pass over the range.  INST: the scene holds instances (without,
// the instance code is compiled out).  SPH: the scene intersects spheres (without, the double-precision sphere
// test is compiled out, which keeps the traversal loop's registers free; no sphere item then occurs).  VOL: the
// scene holds media (without, the free-flight sampling is compiled out; no volume item then occurs).
template <bool INST, bool SPH, bool VOL>
__device__ __forceinline__ bool leaf_test(const SceneView &sv, DRay &ray, int32_t code, int i, float leaf_t,
                                          Best &b, LaneStack &st, const dm::Key &key, uint32_t pixel,
                                          uint32_t sample, uint32_t bounce) {
    const int32_t start = (int32_t)((uint32_t)code & LEAF_CODE_START);
    if (!((uint32_t)code & LEAF_CODE_ITEMS)) {
        const GTri g = sv.tris[start + i];
        float t, u, v, det;
        if (!tri_t(g, ray, T_MIN, b.t, &t, &u, &v, &det)) return false;
        if (!(t < b.t || g.rank > b.rank)) return false;
        if (g.flags & 1) {
            f3 nn;
            if (!tri_normal(sv.tri_normals[start + i], u, v, &nn)) return false;
        }
        b = Best{t, g.rank, make_ref(ITEM_TRIANGLE, start + i), 0u, 0u};
        return false;
    }
    const GItem it = sv.items[start + i];
    if (it.kind == ITEM_LINK) {
        st.push(it.index, leaf_t);
        return false;
    }
    // An object with a non-identity transform (SimpleObject / VolumetricObject::full_intersect around
    // ObjectTransform, object/transform.rs): its primitive or volume is tested against the object-space ray with
    // the interval [T_MIN, +inf), and the hit's distance is that of its mapped position from the world ray's
    // origin (outgoing_intersection).  Testing the object-space distance against the open interval, instead of
    // against the current closest world distance, keeps the closest hit independent of the order objects are
    // tested in.  The hit keeps the object-space distance in its aux word.
    int kind = it.kind;
    int32_t j = it.index;
    DRay r = ray;
    float t_end = b.t;
    const GXform *x = nullptr;
    if (kind == ITEM_XFORM) {
        x = sv.xforms + j;
        if (INST && x->inner_kind == ITEM_LINK) {
            enter_instance(x, start + i, code, i, leaf_t, ray, b, st);
            return true;
        }
        kind = x->inner_kind;
        j = x->inner_index;
        r = object_ray(x, ray);
        t_end = __builtin_inff();
    }
    float t = 0.0f;
    bool ok = false;
    switch (kind) {
        case ITEM_SPHERE: ok = SPH && sphere_t(sv.spheres[j], r, T_MIN, t_end, &t); break;
        case ITEM_TRIANGLE: {
            const GTri g = sv.tris[j];
            float u, v, det;
            ok = tri_t(g, r, T_MIN, t_end, &t, &u, &v, &det);
            if (ok && (g.flags & 1)) {
                f3 nn;
                ok = tri_normal(sv.tri_normals[j], u, v, &nn);
            }
            break;
        }
        case ITEM_PARALLELOGRAM: ok = para_t(sv.paras[j], r, T_MIN, t_end, &t); break;
        case ITEM_AXIS_BOX: {
            int axis;
            float w;
            f3 sg;
            ok = box_t(sv.boxes[j], r, T_MIN, t_end, &t, &axis, &w, &sg);
            break;
        }
        case ITEM_VOLUME: ok = VOL && volume_t<SPH>(sv, sv.volumes[j], r, t_end, key, pixel, sample, bounce, &t); break;
        default: break;
    }
    if (!ok) return false;
    uint32_t aux = 0u;
    if (x) {
        aux = __float_as_uint(t);
        t = length(affine_point(x->fwd, at(r, t)) - ray.o);
        if (!(t <= b.t)) return false;
    }
    if (t < b.t || it.rank > b.rank)
        b = Best{t, it.rank,
                 it.kind == ITEM_TRIANGLE ? make_ref(ITEM_TRIANGLE, it.index) : make_ref(it.kind, start + i), aux,
                 0u};
    return false;
}

// Closest bounded hit of every live path (Renderer::calculate_intersection, renderer.rs:494-501); the
// unbounded objects are marched afterwards by k_march.  BATCH: the slots belong to the requests of bv, each
// with its own scene root and RNG key.  INST: the scene holds instances.  SPH: it intersects spheres.  VOL: it
// holds media.  Occupancy: a 128-lane block holds 11 KiB of LDS stack, so 14 blocks fit a CU.  Without
// instances, spheres or media the kernel runs at seven waves per SIMD (72 VGPRs, a few bytes of spills): on the
// 1080p mesh frame 32.8 ms of k_extend with everything compiled in at six waves, 31.7 without the media code,
// 31.1 at seven waves with 11 LDS entries (31.4 with 10).  The other forms keep six waves (80 VGPRs; at six
// waves with a 16-entry stack against five: 8.5% faster).
// The box loop of k_extend ends for the wave once this many lanes or fewer still search: waiting for the last
// few searching lanes left most of the wave idle (on the 1080p mesh frame a quarter of the loop's lane slots
// belonged to lanes whose traversal had ended and a fifth to lanes stalled on a second range).  Measured there:
// 0 (wait for every lane) 40.0 ms of k_extend per frame, 4: 34.8, 6: 34.4, 8: 34.1, 10: 34.2, 12: 34.5, 16: 35.3.
constexpr uint32_t EXTEND_INNER_EXIT = 8;
// Likewise the primitive loop ends once this many lanes or fewer still test primitives (ranges hold 1 to 8 of
// them, and lanes with short ones waited for the longest).  Measured there with EXTEND_INNER_EXIT 8: 0 (every
// range tested to its end) 34.4-34.6 ms, 4: 33.5, 8: 32.8, 16: 32.2, 24: 32.5, 32: 33.0, 64 (one primitive per
// pass) 33.6; with EXTEND_INNER_EXIT 4 or 16 and this at 16: 32.9 and 33.1.  Holding part-tested ranges of items
// back from speculation (below) left it at 32.6.
constexpr uint32_t EXTEND_PRIM_EXIT = 16;

#ifndef RAYNA_EXT_WAVES
#define RAYNA_EXT_WAVES 6
#endif
#ifndef RAYNA_EXT_WAVES_LEAN
#define RAYNA_EXT_WAVES_LEAN 7
#endif
template <bool BATCH, bool INST, bool SPH, bool VOL>
__global__ __launch_bounds__(EXT_BLOCK) __attribute__((amdgpu_waves_per_eu(
    (INST || SPH || VOL) ? RAYNA_EXT_WAVES : RAYNA_EXT_WAVES_LEAN))) void k_extend(
    SceneView sv, FrameParams fp, PathBuffers pb, const int32_t *queue, const uint32_t *count_ptr, uint32_t depth,
    int2 *overflow, uint32_t *work, ExtendTuning tune, BatchView bv) {
    __shared__ unsigned long long s_entry[EXT_LDS_STACK * EXT_BLOCK];
    const uint32_t count = queue ? *count_ptr : fp.n_slots;
    const uint32_t refill_lanes = tune.refill_lanes, wave_batch = tune.wave_batch;
    dm::Key key{fp.key0, fp.key1};
    const uint32_t gid = blockIdx.x * EXT_BLOCK + threadIdx.x;
    const uint32_t lane = __lane_id();
    LaneStack st{(lds_u64 *)s_entry, overflow + gid, gridDim.x * EXT_BLOCK, 0};
    bool active = false, exhausted = false;
    uint32_t slot = 0, pixel = 0, sample = 0;
    DRay ray{};
    Best b{FLT_MAX, -1, 0u, 0u};
    int32_t leaf = 0;
    int leaf_i = 0, leaf_n = 0;
    float leaf_t = 0.0f;
    // The wave's unhanded queue positions [batch_next, batch_end): claimed wave_batch at a time with one atomic,
    // handed to idle lanes without further atomics.  (All of these are wave-uniform.)  The queue is cut into
    // EXTEND_PARTS contiguous parts with a counter each; a wave starts on part blockIdx.x % EXTEND_PARTS and moves
    // to the next part when its own is used up.  Workgroups go to the eight XCDs in turn, so each part (a band
    // of neighbouring pixels) is worked through by two XCDs, whose L2s then hold its nodes, and no single
    // counter takes every claim of the launch (four parts measured 1% faster than one or eight).
    uint32_t batch_next = 0, batch_end = 0;
    uint32_t part = blockIdx.x % EXTEND_PARTS, parts_done = 0;
    auto part_bound = [&](uint32_t p) { return (uint32_t)(((uint64_t)count * p) / EXTEND_PARTS); };
    RAYNA_STAT_DECL();
    bool last_claim = false;  // every part is used up: no further claims will yield work
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (!exhausted && (idle == ~0ull || (uint32_t)__popcll(idle) >= refill_lanes)) {
            const uint32_t n = (uint32_t)__popcll(idle);
            const uint32_t r = batch_end - batch_next;  // old positions, handed out first
            uint32_t nb = 0, nb_end = 0;                // new batch [nb, nb_end)
            if (r < n && !last_claim) {
                const int leader = __ffsll((long long)idle) - 1;
                for (;;) {
                    const uint32_t lo = part_bound(part), hi = part_bound(part + 1);
                    uint32_t base = 0;
                    if ((int)lane == leader) base = atomicAdd(work + 32 * part, wave_batch);
                    base = __shfl(base, leader);
                    const bool got = base < hi - lo;
                    if (got) {
                        nb = lo + base;
                        nb_end = (hi - nb > wave_batch) ? nb + wave_batch : hi;
                    }
                    if (!got || nb_end == hi) {  // this part is used up
                        part = part + 1 == EXTEND_PARTS ? 0 : part + 1;
                        if (++parts_done == EXTEND_PARTS) last_claim = true;
                    }
                    if (got || last_claim) break;
                }
            }
            const uint32_t idx = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));  // rank among idle lanes
            uint32_t q = 0xffffffffu;
            if (idx < r) q = batch_next + idx;
            else if (idx - r < nb_end - nb) q = nb + (idx - r);
            if (r >= n) {
                batch_next += n;
            } else {
                const uint32_t from_new = (n - r < nb_end - nb) ? n - r : nb_end - nb;
                batch_next = nb + from_new;
                batch_end = nb_end;
            }
            if (last_claim && batch_next == batch_end) exhausted = true;
            if (!active) {
                if (q != 0xffffffffu) {
                    slot = queue ? (uint32_t)queue[q] : q;
                    ray = ray_from_dir(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]},
                                       f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
                    int32_t root = sv.root;
                    if (BATCH) {
                        const RequestParams &R = bv.reqs[bv.slot_req[slot]];
                        const uint32_t l = slot - R.slot_base;
                        pixel = l / R.n_spp;
                        sample = R.s0 + l % R.n_spp;
                        key = dm::Key{R.key0, R.key1};
                        root = R.root;
                    } else {
                        pixel = slot / fp.n_spp;
                        sample = fp.s0 + slot % fp.n_spp;
                    }
                    b = Best{FLT_MAX, -1, 0u, 0u, 0u};
                    st.sp = 0;
                    st.top_valid = false;
                    if (root >= 0) st.push(root, -__builtin_inff());
                    leaf = 0;
                    active = true;
                }
            }
        }
        if (!__any(active)) {
            if (exhausted) break;
            continue;
        }
        RAYNA_STAT(iterations, 1);
        if (active) {
            // Inner nodes until at most EXTEND_INNER_EXIT lanes still search for a leaf range while some lane
            // holds one (or until none searches).  A lane that holds a range keeps expanding inner nodes while
            // others search (speculative traversal, Aila & Laine 2009); a second leaf range, or an instance's
            // end, goes back on the stack and the lane waits.  The lanes still searching when the loop ends keep
            // their stacks and go on in the next pass, after the ranges are tested; every pass tests a range or
            // pops an entry, so the loop ends.
            bool stalled = leaf != 0 && ((uint32_t)leaf & LEAF_CODE_ITEMS) && leaf_i > 0;
            for (;;) {
                const unsigned long long searching = __ballot(leaf == 0 && st.sp > 0);
                if (searching == 0ull) break;
                if ((uint32_t)__popcll(searching) <= EXTEND_INNER_EXIT && __any(leaf != 0)) break;
                if (stalled || st.sp == 0) continue;
                int32_t code;
                float tent;
                st.pop(&code, &tent);
                RAYNA_STAT(pops, 1);
                if (!(tent <= b.t * BOX_SLACK)) {  // culled, or an instance's marker (a NaN distance)
                    if (INST && tent != tent) {
                        if (leaf != 0) {
                            st.push(code, tent);
                            stalled = true;
                        } else {
                            finish_instance(sv, pb, slot, code, ray, b, st);
                        }
                    }
                    continue;
                }
                if (code < 0) {
                    if (leaf != 0) {
                        st.push(code, tent);
                        stalled = true;
                        continue;
                    }
                    leaf = code;
                    leaf_i = 0;
                    leaf_n = (int)(((uint32_t)code >> 27) & 7u) + 1;
                    leaf_t = tent;
                    continue;
                }
                RAYNA_STAT(expands, 1);
                expand(sv, ray, b.t, code, st);
            }
            // The held ranges' primitives, one per pass, until at most EXTEND_PRIM_EXIT lanes still test.  A lane
            // left holding part of a range keeps it (its next entry in leaf_i) through the next box loop: with a
            // range of triangles, which push nothing, it speculates there as a lane holding a range does; with a
            // range of items it waits (stalled), because the links and instance entries its items push count
            // towards the stack bound only as a block on top of a depth-first state, not interleaved with
            // speculative expansions of the links pushed so far.  Every pass tests a primitive of each holder, so
            // the loop ends.
            do {
                if (leaf != 0) {
                    RAYNA_STAT(prims, 1);
                    if (leaf_test<INST, SPH, VOL>(sv, ray, leaf, leaf_i, leaf_t, b, st, key, pixel, sample, depth) ||
                        ++leaf_i == leaf_n)
                        leaf = 0;
                }
            } while ((uint32_t)__popcll(__ballot(leaf != 0)) > EXTEND_PRIM_EXIT);
            if (st.sp == 0 && leaf == 0) {
                pb.hit_t[slot] = b.t;
                pb.hit_ref[slot] = b.ref;
                pb.hit_aux[slot] = b.aux;
                if (INST && b.sub) pb.hit_sub[slot] = b.sub;
                if (pb.hit_rank) pb.hit_rank[slot] = b.rank;
                active = false;
            }
        }
    }
    RAYNA_STAT_FLUSH();
}

// Rays traced as wave packets: at depth 0 of a single-request frame the 64 lanes of a wave hold neighbouring
// samples of one pixel (slot = pixel * n_spp + sample), whose rays start together and nearly agree in direction;
// at the next depths (queue: the paths k_shade kept, in slot order within each strip) a wave's 64 queue positions
// hold the surviving samples of one or two pixels, whose first-bounce rays start from nearly one point (traced
// as packets only on request: packet_depths).  The wave walks the tree once, as a unit: it keeps one stack in
// LDS, whose entries hold an entry
// code, the lanes that take it and the least entry distance among those lanes; node and primitive addresses are
// wave-uniform.  Each lane still tests every box against its own ray and its own closest hit, and every
// primitive of a visited leaf range against its own ray (leaf_test), so a lane visits a superset of the nodes
// its own traversal would and, culling being conservative, its closest hit is the same.  In a batch (BATCH) a
// wave's lanes can belong to requests on different scenes: the wave walks once per scene root among its lanes,
// each walk with the lanes of that root.  Scenes with instances, and trees that need more than PACKET_STACK
// entries, take k_extend.
constexpr int PACKET_STACK = 64;
#ifndef RAYNA_PACKET_MIN_SPP
#define RAYNA_PACKET_MIN_SPP 16
#endif
constexpr uint32_t PACKET_MIN_SPP = RAYNA_PACKET_MIN_SPP;  // fewer samples per pixel spread a wave over pixels

// Least entry distance over the 64 lanes (every lane active), as a key: negative distances (the origin inside
// the box) count as 0, which orders and culls like them (every closest hit lies beyond T_MIN > 0), so the keys
// are non-negative floats, ordered as their bit patterns: an unsigned minimum within rows of 16 by DPP, then over
// the four rows.  (A lane outside the exec mask would read as 0 through bound_ctrl: still a lower bound.)
__device__ __forceinline__ float wave_min_key(float v) {
    uint32_t x = (uint32_t)max(__float_as_int(v), 0);
    x = min(x, (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xf, 0xf, true));
    x = min(x, (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xf, 0xf, true));
    x = min(x, (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xf, 0xf, true));
    x = min(x, (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xf, 0xf, true));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)x, 0), r1 = (uint32_t)__builtin_amdgcn_readlane((int)x, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)x, 32), r3 = (uint32_t)__builtin_amdgcn_readlane((int)x, 48);
    return __uint_as_float(min(min(r0, r1), min(r2, r3)));
}

// SPH and VOL as for k_extend: the sphere test and the media code are compiled in only for scenes that need them.
template <bool BATCH, bool SPH, bool VOL>
__global__ __launch_bounds__(EXT_BLOCK) void k_extend_packet(SceneView sv, FrameParams fp, PathBuffers pb,
                                                             const int32_t *queue, const uint32_t *count_ptr,
                                                             uint32_t depth, BatchView bv) {
    __shared__ int32_t s_code[EXT_BLOCK / 64][PACKET_STACK];
    __shared__ float s_key[EXT_BLOCK / 64][PACKET_STACK];
    __shared__ unsigned long long s_mask[EXT_BLOCK / 64][PACKET_STACK];
    const uint32_t count = queue ? *count_ptr : fp.n_slots;
    if (blockIdx.x * EXT_BLOCK >= count) return;  // the whole block is past the queue's end
    const uint32_t w = threadIdx.x / 64, lane = __lane_id();
    const uint32_t q = blockIdx.x * EXT_BLOCK + threadIdx.x;
    const bool valid = q < count;
    const uint32_t slot = valid ? (queue ? (uint32_t)queue[q] : q) : 0u;
    const uint32_t ls = slot;
    uint32_t pixel, sample;
    dm::Key key{fp.key0, fp.key1};
    int32_t root = sv.root;
    DRay ray;
    const bool computed = !queue && !fp.rays_stored;  // the camera ray, computed as k_generate computes it
    if (BATCH) {
        const RequestParams &R = bv.reqs[bv.slot_req[ls]];
        const uint32_t l = ls - R.slot_base;
        pixel = l / R.n_spp;
        sample = R.s0 + l % R.n_spp;
        key = dm::Key{R.key0, R.key1};
        root = R.root;
        if (computed) {
            const DRay c = camera_ray(R.vp, R.width, R.height, R.sdim, key, pixel, sample);
            ray = ray_from_dir(c.o, c.d);
        }
    } else {
        pixel = ls / fp.n_spp;
        sample = fp.s0 + ls % fp.n_spp;
        if (computed) {
            const DRay c = camera_ray(fp.vp, fp.width, fp.height, fp.sdim, key, pixel, sample);
            ray = ray_from_dir(c.o, c.d);
        }
    }
    if (!computed) ray = ray_from_dir(f3{pb.ox[ls], pb.oy[ls], pb.oz[ls]}, f3{pb.dx[ls], pb.dy[ls], pb.dz[ls]});
    Best b{FLT_MAX, -1, 0u, 0u, 0u};
    // The slab distances as fused multiply-adds, fma(b, 1/d, -o/d): the rounding of -o/d adds an error of at most
    // 2^-24 |o/d| per axis, not relative to t, so the entry distance is lowered and the exit distance raised by
    // 2^-20 max |o/d|, which exceeds it and the roundings of that adjustment, keeping both within the bounds of
    // the two-rounding form (b - o) * (1/d), conservative under the same BOX_SLACK.  A wave with a lane whose ray
    // has a zero direction component (1/d infinite), or |o/d| beyond 2^126, keeps the two-rounding form: there a
    // slab distance is +-inf, or NaN where the ray lies in a box's plane, and the engine's test culls by exactly
    // those values.
    const f3 noi{-(ray.o.x * ray.inv.x), -(ray.o.y * ray.inv.y), -(ray.o.z * ray.inv.z)};
    const float oi_max = fmaxf(fmaxf(fabsf(noi.x), fabsf(noi.y)), fabsf(noi.z));
    const float slack = oi_max * 0x1p-20f;
    // per axis, since the maximum passes over a NaN (0 * inf where the origin lies on the plane o = 0)
    const bool lane_fused = fabsf(noi.x) <= 0x1p126f && fabsf(noi.y) <= 0x1p126f && fabsf(noi.z) <= 0x1p126f;
    const bool fused = !__any(valid && !lane_fused);
    LaneStack unused{nullptr, nullptr, 0, 0};  // leaf_test<false, ...> pushes only for mesh-BVH links, handled here
    int sp = 0;
    auto push = [&](int32_t code, float k, unsigned long long m) {
        s_code[w][sp] = code;
        s_key[w][sp] = k;
        s_mask[w][sp] = m;
        ++sp;
    };
    unsigned long long todo = __ballot(valid);
    while (todo != 0ull) {  // one walk per scene root among the wave's lanes (a single one outside batches)
        int32_t groot = sv.root;
        unsigned long long group = todo;
        if (BATCH) {
            const int leader = __ffsll((long long)todo) - 1;
            groot = __builtin_amdgcn_readlane(root, leader);
            group = __ballot(((todo >> lane) & 1ull) != 0ull && root == groot);
        }
        todo &= ~group;
        if (groot >= 0) push(groot, -__builtin_inff(), group);
        while (sp > 0) {
            --sp;
            const int32_t code = __builtin_amdgcn_readfirstlane(s_code[w][sp]);
            const float kt = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(s_key[w][sp])));
            const unsigned long long m = s_mask[w][sp];
            const bool take = ((m >> lane) & 1ull) != 0ull && kt <= b.t * BOX_SLACK;
            const unsigned long long tm = __ballot(take);
            if (tm == 0ull) continue;
            if (code >= 0) {  // inner node: each child the lanes' boxes hit, far to near
                const float4 *np = reinterpret_cast<const float4 *>(sv.nodes + code);
                const float4 lx = np[0], ly = np[1], lz = np[2], hx = np[3], hy = np[4], hz = np[5];
                const int4 cc = reinterpret_cast<const int4 *>(np)[7];
                const float lo_x[4] = {lx.x, lx.y, lx.z, lx.w}, lo_y[4] = {ly.x, ly.y, ly.z, ly.w},
                            lo_z[4] = {lz.x, lz.y, lz.z, lz.w};
                const float hi_x[4] = {hx.x, hx.y, hx.z, hx.w}, hi_y[4] = {hy.x, hy.y, hy.z, hy.w},
                            hi_z[4] = {hz.x, hz.y, hz.z, hz.w};
                int32_t ccode[4] = {cc.x, cc.y, cc.z, cc.w};
                float ckey[4];
                unsigned long long cmask[4];
    #pragma unroll
                for (int k = 0; k < 4; ++k) {
                    float t0, t1;
                    if (fused) {
                        const float tx1 = fmaf(lo_x[k], ray.inv.x, noi.x), tx2 = fmaf(hi_x[k], ray.inv.x, noi.x);
                        const float ty1 = fmaf(lo_y[k], ray.inv.y, noi.y), ty2 = fmaf(hi_y[k], ray.inv.y, noi.y);
                        const float tz1 = fmaf(lo_z[k], ray.inv.z, noi.z), tz2 = fmaf(hi_z[k], ray.inv.z, noi.z);
                        t0 = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2)) - slack;
                        t1 = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2)) + slack;
                    } else {
                        const float tx1 = (lo_x[k] - ray.o.x) * ray.inv.x, tx2 = (hi_x[k] - ray.o.x) * ray.inv.x;
                        t0 = fminf(tx1, tx2), t1 = fmaxf(tx1, tx2);
                        const float ty1 = (lo_y[k] - ray.o.y) * ray.inv.y, ty2 = (hi_y[k] - ray.o.y) * ray.inv.y;
                        t0 = fmaxf(t0, fminf(ty1, ty2));
                        t1 = fminf(t1, fmaxf(ty1, ty2));
                        const float tz1 = (lo_z[k] - ray.o.z) * ray.inv.z, tz2 = (hi_z[k] - ray.o.z) * ray.inv.z;
                        t0 = fmaxf(t0, fminf(tz1, tz2));
                        t1 = fminf(t1, fmaxf(tz1, tz2));
                    }
                    const float lo = fmaxf(t0, T_MIN), hi = fminf(t1, b.t);
                    const bool hit = take && ccode[k] != ENTRY_NONE && lo <= hi * BOX_SLACK;
                    cmask[k] = __ballot(hit);
                    ckey[k] = cmask[k] != 0ull ? wave_min_key(hit ? t0 : __builtin_inff()) : -__builtin_inff();
                }
    #define RAYNA_PSWAP(a, b)                          \
        if (ckey[a] < ckey[b]) {                       \
            const float tt_ = ckey[a];                 \
            ckey[a] = ckey[b];                         \
            ckey[b] = tt_;                             \
            const int32_t cc_ = ccode[a];              \
            ccode[a] = ccode[b];                       \
            ccode[b] = cc_;                            \
            const unsigned long long mm_ = cmask[a];   \
            cmask[a] = cmask[b];                       \
            cmask[b] = mm_;                            \
        }
                RAYNA_PSWAP(0, 1)
                RAYNA_PSWAP(2, 3)
                RAYNA_PSWAP(0, 2)
                RAYNA_PSWAP(1, 3)
                RAYNA_PSWAP(1, 2)
    #undef RAYNA_PSWAP
    #pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (cmask[k] != 0ull) push(ccode[k], ckey[k], cmask[k]);
                continue;
            }
            // leaf range: mesh-BVH links go on the wave's stack for the lanes that took the range; the other
            // entries are tested by those lanes
            const int n = (int)(((uint32_t)code >> 27) & 7u) + 1;
            const int32_t start = (int32_t)((uint32_t)code & LEAF_CODE_START);
            const bool items = ((uint32_t)code & LEAF_CODE_ITEMS) != 0u;
            for (int i = 0; i < n; ++i) {
                if (items && sv.items[start + i].kind == ITEM_LINK) {
                    push(sv.items[start + i].index, kt, tm);
                    continue;
                }
                if (take) (void)leaf_test<false, SPH, VOL>(sv, ray, code, i, kt, b, unused, key, pixel, sample, depth);
            }
        }
    }
    if (valid) {
        pb.hit_t[slot] = b.t;
        pb.hit_ref[slot] = b.ref;
        pb.hit_aux[slot] = b.aux;
        if (pb.hit_rank) pb.hit_rank[slot] = b.rank;
    }
}
