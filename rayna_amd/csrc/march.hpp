// The unbounded SDF objects of a closest-hit query, marched apart from the tree (included by kernels.hip inside
// its anonymous namespace, after sdf_eval).
//
// Work item = (path, unbounded object), SDF-major (item = j * count + q), so the items a wave takes mostly
// belong to one SDF.  Each pass of the persistent loop takes every active lane one march step: cut and escape
// tests, the estimate (a Mandelbulb estimate's iterations complete within the pass; lanes whose estimate bails out
// early wait for the wave's longest, which costs less than handing the wave back to step starts and ends between
// iterations), and the end of the step.  A lane whose march ends takes the next item of its wave's claimed batch
// in the next pass.  The operations of each march are exactly those of RaymarchedIsosurfaceMesh::intersect
// (raymarched.rs:73-122) and sdf_eval, in the same order, so the outcome is the same bit for bit: the march
// starts at total = T_MIN; each step evaluates the estimate at the ray point of `total`, advances total by its
// magnitude and hits when that magnitude is below epsilon, giving up after max_iterations + 2 estimates.  Two
// exits end a march early with the same outcome as the full march: the march distance never decreases, so once
// it reaches the cut no later hit can be kept; and once the point is outside the object's escape sphere and the
// ray moves away from its centre, every later estimate exceeds epsilon.
//
// Each item's outcome is written to mt[item]: +t (front side), -t (back side) or +infinity (no hit), and at a hit
// the march iteration to mt[mt_stride + item] (as integer bits; the hit's side).  The
// cut is the path's closest bounded hit; k_march_merge then applies the list order (strict <, object/list.rs:
// 358-363).  A cut that ignores earlier objects of the list only lets a march run longer: it never changes which
// hit wins.

constexpr int MARCH_BLOCK = 256;
// A wave's march claim takes this many times the extend kernels' claim (tuning's wave_batch) from the one shared
// counter: with claims as small as the extend kernels' the waves waited on that counter (the SDF frame at 128,
// 512, 1024 and 2048 items per claim for both kinds of kernel: 85.6, 76.7, 79.3 and 89.3 ms).
constexpr uint32_t MARCH_CLAIM_FACTOR = 4;

// The state of a lane's march (k_march): its item, ray and cut, the march distance and step.
struct MarchLane {
    uint32_t item;
    DRay ray;
    float t_cut, total;
    int step;
};

// One step of a lane's march against SDF f: the cut and escape tests, the estimate at the ray point of `total`, and
// the step's end.  False once the march has ended (its outcome written to mt).
__device__ __forceinline__ bool march_step(const GSdf &f, const GSphere *spheres, MarchLane &L, float *mt,
                                           uint64_t mt_stride) {
    const f3 point = at(L.ray, L.total);
    const f3 qe = point - f3{f.ex, f.ey, f.ez};
    if (L.total >= L.t_cut || (dot(qe, qe) > f.escape_r * f.escape_r && dot(qe, L.ray.d) > 0.0f)) {
        mt[L.item] = __builtin_inff();
        return false;
    }
    const float dist = sdf_eval(spheres, f, point);
    L.total = L.total + fabsf(dist);
    if (fabsf(dist) < f.epsilon && T_MIN <= L.total && L.total <= FLT_MAX) {
        mt[L.item] = signbit(dist) ? -L.total : L.total;
        mt[mt_stride + L.item] = __int_as_float(L.step);  // Intersection::side (raymarched.rs:109)
        return false;
    }
    if (L.step > f.max_iterations) {
        mt[L.item] = __builtin_inff();
        return false;
    }
    ++L.step;
    return true;
}

template <bool BATCH>
__global__ __launch_bounds__(MARCH_BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_march(
    FrameParams fp, PathBuffers pb, const GSdf *__restrict__ sdfs, const GSphere *__restrict__ spheres,
    const int32_t *queue, const uint32_t *count_ptr, uint32_t n_sdf, float *mt, uint64_t mt_stride, uint32_t *work,
    ExtendTuning tune, BatchView bv) {
    const uint32_t count = queue ? *count_ptr : fp.n_slots;
    const uint32_t n_items = count * n_sdf;
    const uint32_t refill_lanes = tune.refill_lanes, wave_batch = tune.wave_batch * MARCH_CLAIM_FACTOR;
    const uint32_t lane = __lane_id();
    bool active = false, exhausted = false, last_claim = false;
    uint32_t batch_next = 0, batch_end = 0;
    int32_t j = 0;  // the lane's SDF
    MarchLane L{};
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        if (!exhausted && idle != 0ull) {
            const uint32_t n = (uint32_t)__popcll(idle);
            const uint32_t rem = batch_end - batch_next;
            uint32_t nb = 0, nb_end = 0;
            if (rem < n && !last_claim && (idle == ~0ull || n >= refill_lanes)) {
                const int leader = __ffsll((long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(work, wave_batch);
                base = __shfl(base, leader);
                nb = base < n_items ? base : n_items;
                nb_end = (n_items - nb > wave_batch) ? nb + wave_batch : n_items;
                if (nb_end >= n_items) last_claim = true;
            }
            const uint32_t idx = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
            uint32_t q = 0xffffffffu;
            if (idx < rem) q = batch_next + idx;
            else if (idx - rem < nb_end - nb) q = nb + (idx - rem);
            if (rem >= n) {
                batch_next += n;
            } else if (nb_end > nb) {
                const uint32_t from_new = (n - rem < nb_end - nb) ? n - rem : nb_end - nb;
                batch_next = nb + from_new;
                batch_end = nb_end;
            } else {
                batch_next = batch_end;
            }
            if (last_claim && batch_next == batch_end) exhausted = true;
            if (!active && q != 0xffffffffu) {
                const uint32_t jj = q / count, qq = q % count;
                const uint32_t slot = queue ? (uint32_t)queue[qq] : qq;
                int32_t first = 0, n_own = (int32_t)n_sdf;
                if (BATCH) {
                    const RequestParams &R = bv.reqs[bv.slot_req[slot]];
                    first = R.first_sdf;
                    n_own = R.n_sdfs;
                }
                if ((int32_t)jj >= n_own) {
                    mt[q] = __builtin_inff();  // the request's scene has fewer unbounded objects
                } else {
                    j = first + (int32_t)jj;
                    L.item = q;
                    L.ray = ray_from_dir(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]},
                                         f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
                    L.t_cut = pb.hit_t[slot];
                    L.total = T_MIN;
                    L.step = 0;
                    active = true;
                }
            }
        }
        if (!__any(active)) {
            if (exhausted) break;
            continue;
        }
        if (!active) continue;
        // The items a wave holds normally belong to one SDF: its record is then read at a wave-uniform address of a
        // read-only argument (scalar loads), and the pass's branches on it are uniform.
        const int32_t j0 = __builtin_amdgcn_readfirstlane(j);
        if (__all(j == j0))
            active = march_step(sdfs[j0], spheres, L, mt, mt_stride);
        else
            active = march_step(sdfs[j], spheres, L, mt, mt_stride);
    }
}

// The closest hit of each path over its bounded hit and its unbounded objects' marches, in list order with
// strict < (object/list.rs:120-125).  The paths whose closest hit becomes an unbounded object are appended to
// sdf_queue for k_sdf_normal: each wave takes strips of MERGE_STRIP * 64 queue positions, stages those slots in
// its part of LDS and appends them with one atomic per strip.
constexpr uint32_t MERGE_STRIP = 16;

template <bool BATCH>
__global__ __launch_bounds__(BLOCK) void k_march_merge(SceneView sv, FrameParams fp, PathBuffers pb,
                                                       const int32_t *queue, const uint32_t *count_ptr,
                                                       uint32_t n_sdf, const float *mt, uint64_t mt_stride,
                                                       int32_t *sdf_queue, uint32_t *sdf_count, BatchView bv) {
    const uint32_t count = queue ? *count_ptr : fp.n_slots;
    __shared__ uint32_t staged[BLOCK * MERGE_STRIP];
    uint32_t *mine = staged + (threadIdx.x / 64) * 64 * MERGE_STRIP;
    const uint32_t lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t n_waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64;
    for (uint32_t base = wave * 64 * MERGE_STRIP; base < count; base += n_waves * 64 * MERGE_STRIP) {
      uint32_t n_kept = 0;
      for (uint32_t jx = 0; jx < MERGE_STRIP; ++jx) {
        const uint32_t q = base + jx * 64 + lane;
        bool changed = false;
        uint32_t slot = 0;
        if (q < count) {
            slot = queue ? (uint32_t)queue[q] : q;
            int32_t first = 0, n_own = (int32_t)n_sdf;
            if (BATCH) {
                const RequestParams &R = bv.reqs[bv.slot_req[slot]];
                first = R.first_sdf;
                n_own = R.n_sdfs;
            }
            float bt = pb.hit_t[slot];
            int32_t rank = pb.hit_rank ? pb.hit_rank[slot] : 0;
            uint32_t ref = pb.hit_ref[slot], aux = pb.hit_aux[slot];
            for (int32_t jj = 0; jj < n_own; ++jj) {
                const float v = mt[(size_t)jj * count + q];
                const float t = fabsf(v);
                if (t < bt) {
                    const int32_t j = first + jj;
                    bt = t;
                    rank = -2 - sv.sdfs[j].list_pos;
                    ref = make_ref(ITEM_SDF, j);
                    aux = (signbit(v) ? HIT_AUX_BACK : 0u) |
                          ((uint32_t)__float_as_int(mt[mt_stride + (size_t)jj * count + q]) << HIT_AUX_SIDE_SHIFT);
                    changed = true;
                }
            }
            if (changed) {
                pb.hit_t[slot] = bt;
                pb.hit_ref[slot] = ref;
                pb.hit_aux[slot] = aux;
                if (pb.hit_rank) pb.hit_rank[slot] = rank;
            }
        }
        const unsigned long long mask = __ballot(changed);
        if (changed) mine[n_kept + (uint32_t)__popcll(mask & below)] = slot;
        n_kept += (uint32_t)__popcll(mask);
      }
      if (n_kept != 0u) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t wbase = 0;
        if (lane == 0u) wbase = atomicAdd(sdf_count, n_kept);
        wbase = __shfl(wbase, 0);
        for (uint32_t i = lane; i < n_kept; i += 64) sdf_queue[wbase + i] = (int32_t)mine[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
}

// The shading normal of every path whose closest hit is an unbounded object: sdf_normal (six estimates) at the
// hit point, computed as hit_geometry would compute it, into the path's sdf_n fields.  A dense pass over the
// queue k_march_merge built, so that the estimates run on full waves instead of on the few lanes of each shade
// wave whose path hit an SDF.
__global__ __launch_bounds__(BLOCK) void k_sdf_normal(SceneView sv, PathBuffers pb, const int32_t *sdf_queue,
                                                      const uint32_t *sdf_count) {
    const uint32_t n = *sdf_count;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const uint32_t slot = (uint32_t)sdf_queue[i];
        const DRay ray = ray_lazy(f3{pb.ox[slot], pb.oy[slot], pb.oz[slot]}, f3{pb.dx[slot], pb.dy[slot], pb.dz[slot]});
        const uint32_t ref = pb.hit_ref[slot];
        const f3 nrm = sdf_normal(sv, sv.sdfs[ref & HIT_INDEX_MASK], at(ray, pb.hit_t[slot]));
        pb.sdf_nx[slot] = nrm.x;
        pb.sdf_ny[slot] = nrm.y;
        pb.sdf_nz[slot] = nrm.z;
    }
}
