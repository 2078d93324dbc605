// Interface between the host renderer and the wavefront kernels of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.hpp"

namespace rayna {

// Path state of one chunk of samples, structure-of-arrays in HBM.  A chunk holds samples [s0, s0 + n_spp) of
// every pixel; slot i is sample s0 + i % n_spp of pixel i / n_spp, so a pixel's samples are neighbours (a wave
// traces 64 samples of one pixel, or all the samples of a few neighbouring pixels).
struct PathBuffers {
    float *ox, *oy, *oz;     // ray origin
    float *dx, *dy, *dz;     // ray direction (unit)
    float *tr, *tg, *tb;     // path throughput
    float *lr, *lg, *lb;     // radiance gathered so far; the sample's value once the path ends
    float *hit_t;            // closest-hit distance
    uint32_t *hit_ref;       // kind << 28 | index, 0 on a miss
    uint32_t *hit_aux;       // HIT_AUX_* bits; the object-space distance of a transformed object's hit
    uint32_t *hit_sub;       // an instance's hit: the reference of its primitive inside the instanced mesh BVH
    int32_t *hit_rank;       // optional (parity queries): rank of the closest primitive
    float *sdf_nx, *sdf_ny, *sdf_nz;  // scenes with unbounded objects: the shading normal of an SDF hit
};

struct FrameParams {
    GViewport vp;
    uint32_t width, height;
    uint32_t spp;            // samples per pixel of the frame
    uint32_t sdim;           // isqrt(spp): the stratified grid is sdim x sdim
    uint32_t key0, key1;     // Philox key
    uint32_t ray_depth;
    uint32_t n_pixels;
    uint32_t s0;             // first sample of the chunk
    uint32_t n_spp;          // samples per pixel in the chunk
    uint32_t n_slots;        // slots in the chunk: n_pixels * n_spp
    int32_t mode;            // RAYNA_MODE_*
    uint32_t *checks = nullptr;  // validation on: RAYNA_CHECKS violation counters (RAYNA_CHECK_*), else null
    uint32_t rays_stored = 0;    // k_generate stored this chunk's camera rays (else the first kernels compute them)
};

// One request of a batch, as seen by one chunk of paths: its slots are [slot_base, slot_base + n_slots), sample
// s0 + l % n_spp of pixel l / n_spp for local slot l (pixel-major, like a single frame).
struct RequestParams {
    GViewport vp;
    uint32_t width, height, spp, sdim;
    uint32_t key0, key1, ray_depth, n_pixels;
    uint32_t s0, n_spp, slot_base, n_slots;
    int32_t root, first_sdf, n_sdfs, skybox;  // the request's scene within the merged scene set
    int32_t sky_image;                        // RAYNA_SKY_HDRI: its image
    uint32_t sum_offset;                      // its 3 * n_pixels floats in the batch's sum and output buffers
    int32_t layout;                           // output layout (RAYNA_LAYOUT_*)
    uint32_t first_piece;                     // 1 if the chunk holds the request's first samples
};

// The requests of a batch chunk.  reqs == nullptr: a single frame described by FrameParams and SceneView.
struct BatchView {
    const RequestParams *reqs;
    uint32_t n_reqs;
    uint32_t *slot_req;  // request (index into reqs) of every slot, written by generate
};

// Inputs and outputs of the adjoint replay (path replay backpropagation) of a frame.
struct AdjointParams {
    const float *d_image;    // dloss/d(frame mean), 3 floats per pixel in row-major pixel order
    double inv_spp;          // 1 / samples per pixel: the share of each sample
    double *grad;            // copies x n_materials x 3 gradient accumulators
    uint32_t n_materials;
    uint32_t copies;         // gradient copies (a block adds into copy blockIdx % copies)
};

// Traversal tuning of the extend kernel: a wave takes new rays once `refill_lanes` of its lanes are idle,
// claiming `wave_batch` queue positions per atomic.
// packet_depths: the bounces (depth < packet_depths) traced as wave packets (k_extend_packet) where a chunk's
// samples of a pixel are neighbouring slots.
struct ExtendTuning {
    uint32_t refill_lanes = 32;
    uint32_t wave_batch = 128;
    uint32_t packet_depths = 1;
};

// Shade grid: blocks of 256 threads per compute unit (grid-stride over the queue).
constexpr int SHADE_BLOCKS_PER_CU = 16;

// The camera ray of every slot (the throughput and radiance fields are left as they are: the first shade takes
// them as 1 and 0, and the adjoint replay reads the primal radiance still held there).  With bv.reqs,
// fp.n_slots is the chunk's slot count and the requests come from bv.
void launch_generate(const FrameParams &fp, const PathBuffers &pb, const BatchView &bv, hipStream_t s);
// Claim counters of one extend launch: the queue is cut into EXTEND_PARTS contiguous parts, each with its own
// counter, one 128-byte line apart (EXTEND_CLAIM_WORDS words in all, zeroed before the launch).
#ifndef RAYNA_EXTEND_PARTS
#define RAYNA_EXTEND_PARTS 4
#endif
constexpr uint32_t EXTEND_PARTS = RAYNA_EXTEND_PARTS;
constexpr uint32_t EXTEND_CLAIM_WORDS = EXTEND_PARTS * 32;

// queue == nullptr: identity queue of fp.n_slots entries; else queue[0 .. *count)
// overflow: traversal stack entries beyond the LDS part, extend_block() * grid * (stack_need - extend_lds_stack());
// work: EXTEND_CLAIM_WORDS zeroed words, the counters the launch's waves take rays from
void launch_extend(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue,
                   const uint32_t *count, uint32_t depth, int grid, int2 *overflow, uint32_t *work,
                   const ExtendTuning &tune, const BatchView &bv, hipStream_t s);
void launch_shade(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue_in,
                  const uint32_t *count_in, int32_t *queue_out, uint32_t *count_out, uint32_t depth, int grid,
                  const AdjointParams *adjoint, const BatchView &bv, hipStream_t s);
// The unbounded SDF objects of each queued path (n_sdf per path; in a batch, each request's own), marched
// after launch_extend with its hit as the cut, then merged into the hit record; the paths whose closest hit
// becomes an SDF get their shading normal in pb.sdf_n*.  mt: two planes of mt_stride >= count * n_sdf floats
// (outcome, march iteration).  work: a zeroed counter.  sdf_queue: room for every queued path; sdf_count: one
// counter (zeroed here).  No-op for n_sdf == 0.
void launch_march(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, const int32_t *queue,
                  const uint32_t *count, uint32_t n_sdf, float *mt, uint64_t mt_stride, uint32_t *work, int grid,
                  int merge_grid, const ExtendTuning &tune, const BatchView &bv, int32_t *sdf_queue,
                  uint32_t *sdf_count, hipStream_t s);
// Pixels per block of the resolve kernels for n samples per pixel in the chunk: a block stages its pixels'
// samples (three channels, rows of n | 1 floats) in LDS so that its loads are coalesced; from 1040 samples per
// pixel on, a block sums 256 pixels straight from HBM.  A small stage keeps many blocks per CU: the 1080p,
// 64-sample resolve took 0.459 ms with 4160 floats per channel, 0.324 with 2080, 0.279 with 1040.
constexpr uint32_t RESOLVE_STAGE = 1040;  // floats per channel
constexpr uint32_t resolve_pixels_per_block(uint32_t n) {
    return (n | 1u) > RESOLVE_STAGE ? 256u : (RESOLVE_STAGE / (n | 1u) > 256u ? 256u : RESOLVE_STAGE / (n | 1u));
}
// Batch chunks: per request, the running per-pixel sum of the chunk's samples (in sample order) and, at the
// request's last chunk, its sample mean in its layout at out + sum_offset.  blocks: the most resolve blocks of
// one request of the chunk (its pixels / resolve_pixels_per_block(its n_spp), rounded up).
void launch_resolve_batch(const BatchView &bv, uint32_t blocks, const PathBuffers &pb, float *sum, hipStream_t s);
void launch_finalize_batch(const BatchView &bv, const uint32_t *final_reqs, uint32_t n_final, uint32_t max_pixels,
                           const float *sum, float *out, hipStream_t s);
void launch_sum_grad(const double *grad, uint32_t n, uint32_t copies, double *out, hipStream_t s);
// The non-PBR render modes: the mode colour of each slot's camera ray (after launch_extend at depth 0) into
// its radiance field.
void launch_shade_mode(const SceneView &sv, const FrameParams &fp, const PathBuffers &pb, int grid, hipStream_t s);
// The UI's 8-bit conversion of a row-major float image (width * height * 3) to RGBA.
void launch_to_srgb8(const float *img, uint32_t n_pixels, uint8_t *rgba, hipStream_t s);
void launch_resolve(const FrameParams &fp, const PathBuffers &pb, float *sum, uint32_t chunk_samples, int first,
                    hipStream_t s);
void launch_finalize(const FrameParams &fp, const float *sum, float *acc, double frames, float *out, int layout,
                     hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t s);
// Whether launch_extend traces this depth as wave packets (k_extend_packet).  At depth 0 of a single request the
// packets compute the camera rays themselves, so only the march and the non-PBR modes need k_generate's rays.
bool extend_packets(const SceneView &sv, const FrameParams &fp, const ExtendTuning &tune, const BatchView &bv,
                    uint32_t depth);
// The VALU issue-rate calibration kernel over `blocks` blocks of 256 lanes (out: one float per lane); returns
// the wave64 VALU instructions its chains issue (loop and set-up instructions not counted).
uint64_t launch_valu_peak(float *out, uint32_t blocks, uint32_t iters, hipStream_t s);
void launch_sqrt_check(uint32_t first, uint32_t count, unsigned long long *mismatches, uint32_t blocks, hipStream_t s);
void launch_count_segments(const uint32_t *queue_count, uint32_t n, uint32_t first_count,
                           unsigned long long *segments, hipStream_t s);

// The static workspace of rayna_adjoint_static (a file-scope device array of adjoint_scratch_doubles()
// doubles): its address on `device`, which must be the current device (taken on the first call per device).
hipError_t adjoint_scratch(int device, double **ptr);
size_t adjoint_scratch_doubles();

int extend_lds_stack();
int extend_blocks_per_cu();  // resident extend blocks per CU (the persistent grid is this times the CU count)
int extend_block();
int march_blocks_per_cu();
int march_block();

}  // namespace rayna
