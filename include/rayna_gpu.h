/*
 * rayna_gpu.h — C ABI of the MI355X (gfx950) renderer for rayna_engine's render hot path.
 *
 * The ABI replaces `rayna_engine::render::renderer::Renderer` (rayna_engine/src/render/renderer.rs:36-264)
 * for the scene graph that the engine's `StandardScene` presets build: a flat list of objects
 * (`Vec<SimpleObject|VolumetricObject>.into()` -> `ObjectList`, rayna_engine/src/object/list.rs:43-129),
 * each a mesh plus a material and an optional `ObjectTransform`, with a `SkyboxInstance`.  Scenes cross the
 * ABI as plain arrays: there are no closures and no `Dynamic*` plug-ins (the renderer returns
 * RAYNA_UNSUPPORTED for what it cannot express).
 *
 * All geometry is given in double precision, as the engine's `Number = f64`
 * (rayna_engine/src/core/types.rs:7); colours are single precision (`Channel = f32`, types.rs:3).
 *
 * Threading mirrors `Renderer::render(&mut self)`: one handle is not thread-safe; separate handles are.
 * The caller owns every host buffer it passes in.
 */
#ifndef RAYNA_GPU_H
#define RAYNA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */

typedef enum rayna_status {
    RAYNA_OK = 0,
    RAYNA_INVALID_ARGUMENT = 1, /* a bad index, size or null pointer in the request */
    RAYNA_UNSUPPORTED = 2,      /* valid for the engine but not for this renderer (e.g. ray_branching > 1) */
    RAYNA_DEVICE_ERROR = 3,     /* a HIP call failed; rayna_last_error() names it */
    RAYNA_OUT_OF_MEMORY = 4,
    RAYNA_RENDERER_LOST = 5     /* a launch on this renderer failed earlier (that call returned RAYNA_DEVICE_ERROR):
                                   its device state is unknown, so every later call is refused; destroy it and create
                                   a new one, as the engine's caller restarts a dead render worker
                                   (rayna_ui/src/app.rs:483-495, integration/mod.rs:100-125) */
} rayna_status;

/* ---------------------------------------------------------------- materials
 * rayna_engine/src/material/mod.rs:168-175 (MaterialInstance).  The albedo (emission for LIGHT) is the solid
 * `colour` (SolidTexture, rayna_engine/src/texture/solid.rs:29) when `texture` is 0, else the texture
 * textures[texture - 1] of the scene.
 *   LAMBERTIAN  lambertian.rs:28-48   colour = albedo
 *   METAL       metal.rs:17-45         colour = albedo, fuzz
 *   DIELECTRIC  dielectric.rs:19-79   colour = albedo, refractive_index, density
 *   ISOTROPIC   isotropic.rs:29-52    colour = albedo, density
 *   LIGHT       light.rs:17-32        colour = emissive
 */
enum {
    RAYNA_MAT_LAMBERTIAN = 0,
    RAYNA_MAT_METAL = 1,
    RAYNA_MAT_DIELECTRIC = 2,
    RAYNA_MAT_ISOTROPIC = 3,
    RAYNA_MAT_LIGHT = 4
};

typedef struct rayna_material {
    int32_t kind;
    float colour[3];
    int32_t texture;  /* 0: the solid colour; k: textures[k - 1] */
    double fuzz;
    double refractive_index;
    double density;
} rayna_material;

/* ---------------------------------------------------------------- meshes
 * rayna_engine/src/mesh/mod.rs:68-85 (MeshInstance).
 *   SPHERE         SphereMesh::new(centre, radius)              primitive/sphere.rs:27-36
 *   TRIANGLE       one Triangle::new(vertices, normals)         primitive/triangle.rs:22-35
 *   TRIANGLE_LIST  MeshList of Triangles (its own SAH BVH)      advanced/list.rs:30-59
 *   PARALLELOGRAM  ParallelogramMesh::new(Planar::new(p,u,v))   planar/parallelogram.rs:25-37
 *   AXIS_BOX       AxisBoxMesh::new(a, b)                       primitive/axis_box.rs:28-37
 *   SDF            RaymarchedIsosurfaceMesh::new_custom         isosurface/raymarched.rs:51-57
 *   BVH            BvhMesh::new(meshes) (its own SAH BVH)        advanced/bvh.rs:28-45
 * `index` selects the element of the kind's array; TRIANGLE_LIST uses triangles [index, index+count); BVH
 * uses the meshes [index, index+count) of the mesh array itself, each a SPHERE, TRIANGLE, PARALLELOGRAM or
 * AXIS_BOX mesh.
 */
enum {
    RAYNA_MESH_SPHERE = 0,
    RAYNA_MESH_TRIANGLE = 1,
    RAYNA_MESH_TRIANGLE_LIST = 2,
    RAYNA_MESH_PARALLELOGRAM = 3,
    RAYNA_MESH_AXIS_BOX = 4,
    RAYNA_MESH_SDF = 5,
    RAYNA_MESH_BVH = 6
};

typedef struct rayna_mesh {
    int32_t kind;
    int32_t index;
    int32_t count;
    int32_t reserved;
} rayna_mesh;

/* ---------------------------------------------------------------- signed distance functions
 * The engine takes any `Fn(Point3) -> Number` (isosurface/mod.rs:7-9); closures cannot cross a C ABI,
 * so the SDF is one of an enumerated set:
 *   SPHERE       |p - c| - r of sphere `first_sphere`
 *   SMIN_SPHERES polynomial smooth-min (parameter k) folded over spheres [first_sphere, +n_spheres)
 *   MANDELBULB   power-`power` Mandelbulb distance estimate about `centre`, `bulb_iterations`
 *                iterations, bail-out radius `bailout`
 * `max_iterations` and `epsilon` are the march limits of RaymarchedIsosurfaceMesh.
 */
enum {
    RAYNA_SDF_SPHERE = 0,
    RAYNA_SDF_SMIN_SPHERES = 1,
    RAYNA_SDF_MANDELBULB = 2
};

typedef struct rayna_sdf {
    int32_t kind;
    int32_t first_sphere;
    int32_t n_spheres;
    int32_t max_iterations;
    int32_t bulb_iterations;
    int32_t reserved;
    double epsilon;
    double k;
    double power;
    double bailout;
    double centre[3];
} rayna_sdf;

/* ---------------------------------------------------------------- objects
 * rayna_engine/src/object/mod.rs:47-52 (ObjectInstance).
 *   SIMPLE      SimpleObject::new_uncorrected(mesh, material, transform)   object/simple.rs:78-92
 *   VOLUMETRIC  VolumetricObject::new_uncorrected(mesh, material, density, transform)
 *                                                                          object/volumetric.rs:42-72
 * `transform` selects the object's ObjectTransform (object/transform.rs): 0 is the identity, k > 0 the matrix
 * transforms[k - 1] of the scene, mapping object space to world space.  The engine's corrected constructors
 * (SimpleObject::new, ObjectTransform::new_corrected) rotate about the mesh's centre; a caller passes the
 * corrected matrix, translate(-centre) then transform then translate(centre).  Non-identity transforms are
 * supported on every bounded mesh: primitives (simple and volumetric objects), and triangle lists and BVH
 * meshes (simple objects: an instance of the mesh's own tree); on SDFs they return RAYNA_UNSUPPORTED.  Volumes
 * take sphere and axis-box boundaries.
 */
enum {
    RAYNA_OBJ_SIMPLE = 0,
    RAYNA_OBJ_VOLUMETRIC = 1
};

typedef struct rayna_object {
    int32_t kind;
    int32_t mesh;
    int32_t material;
    int32_t transform;  /* 0: identity; k: transforms[k - 1] */
    double density;
} rayna_object;

/* ---------------------------------------------------------------- textures
 * rayna_engine/src/texture/mod.rs:26-37 (TextureInstance):
 *   SOLID          SolidTexture { albedo = colour }                        texture/solid.rs:8-30
 *   WORLD_CHECKER  WorldCheckerTexture { offset xyz, even, odd, scale }    texture/checker.rs:12-26
 *   UV_CHECKER     UvCheckerTexture { offset xy, even, odd, scale }        texture/checker.rs:28-42
 *   IMAGE          ImageTexture { image, scale = image_scale, offset xy }  texture/image.rs:7-40
 *   UV_NOISE       UvNoiseTexture, the noise at (u, v, 0)                  texture/noise.rs:62-70
 *   WORLD_NOISE    WorldNoiseTexture, the noise at the hit position        texture/noise.rs:84-92
 *   LOCAL_NOISE    LocalNoiseTexture, the noise at the hit's local position texture/noise.rs:106-114
 * `odd` and `even` name textures as a material does (k: textures[k - 1]) and must name earlier textures.
 * The noise textures take ScalePoint(noise).set_scale(scale) of an improved Perlin noise (K. Perlin, "Improving
 * Noise", SIGGRAPH 2002) whose permutation is drawn from `noise_seed[i]`; the engine's `noise` crate generates
 * its own permutation, so the values differ from the engine's for the same seed.  `noise_rgb` = 0 is
 * ColourSource::Greyscale (seed 0), 1 ColourSource::Rgb (seeds 0, 1, 2); both map [-1, 1] to [0, 1].
 */
enum {
    RAYNA_TEX_SOLID = 0,
    RAYNA_TEX_WORLD_CHECKER = 1,
    RAYNA_TEX_UV_CHECKER = 2,
    RAYNA_TEX_IMAGE = 3,
    RAYNA_TEX_UV_NOISE = 4,
    RAYNA_TEX_WORLD_NOISE = 5,
    RAYNA_TEX_LOCAL_NOISE = 6
};

typedef struct rayna_texture {
    int32_t kind;
    float colour[3];
    int32_t odd, even;     /* checkers */
    int32_t image;         /* IMAGE: index into the scene's images */
    int32_t noise_rgb;     /* noise: 0 greyscale, 1 rgb */
    uint32_t noise_seed[3];
    int32_t reserved;
    double scale;          /* checkers: cell size; noise: point scale */
    double offset[3];      /* WORLD_CHECKER xyz; UV_CHECKER and IMAGE xy */
    double image_scale[2]; /* IMAGE: uv scale */
} rayna_texture;

/* An image (core/image.rs): width x height RGB floats, row-major (pixel (x, y) at rgb[3 * (y * width + x)]),
 * sampled bilinearly with clamped corners as Image::get_bilinear does (core/image.rs:86-110). */
typedef struct rayna_image {
    uint32_t width, height;
    const float *rgb;
} rayna_image;

/* ---------------------------------------------------------------- skybox (skybox/mod.rs:26-43) */
enum {
    RAYNA_SKY_NONE = 0,   /* NoSkybox       skybox/none.rs:8-13 */
    RAYNA_SKY_SIMPLE = 1, /* SimpleSkybox   skybox/simple.rs:12-22 */
    RAYNA_SKY_WHITE = 2,  /* WhiteSkybox    skybox/simple.rs:26-30 */
    RAYNA_SKY_HDRI = 3    /* HdrImageSkybox skybox/hdri.rs:8-27: images[sky_image] */
};

/* The scene: the object list in insertion order (the order `ObjectList::new_uncorrected` sees), and
 * flat parameter arrays.  spheres [n][4] = centre xyz, radius; tri_vertices and tri_normals [n][3][3];
 * parallelograms [n][9] = p, u, v; boxes [n][6] = corner a, corner b; transforms [n][12] = the affine matrix
 * row by row, [m00 m01 m02 t0; m10 m11 m12 t1; m20 m21 m22 t2] (point p maps to M p + t). */
typedef struct rayna_scene_desc {
    const rayna_object *objects;
    int32_t n_objects;
    const rayna_mesh *meshes;
    int32_t n_meshes;
    const rayna_material *materials;
    int32_t n_materials;
    const double *spheres;
    int32_t n_spheres;
    const double *tri_vertices;
    const double *tri_normals;
    int32_t n_triangles;
    const double *parallelograms;
    int32_t n_parallelograms;
    const double *boxes;
    int32_t n_boxes;
    const rayna_sdf *sdfs;
    int32_t n_sdfs;
    int32_t skybox;
    const double *transforms;
    int32_t n_transforms;
    const rayna_texture *textures;
    int32_t n_textures;
    const rayna_image *images;
    int32_t n_images;
    int32_t sky_image;     /* RAYNA_SKY_HDRI: the sky's image */
} rayna_scene_desc;

/* ---------------------------------------------------------------- camera (scene/camera.rs:10-38); angles in radians */
typedef struct rayna_camera {
    double pos[3];
    double fwd[3];
    double v_fov;
    double focus_dist;
    double defocus_angle;
} rayna_camera;

/* ---------------------------------------------------------------- render options (render/render_opts.rs:8-29)
 * `seed` keys the counter-based RNG that replaces the engine's entropy-seeded generators
 * (render/renderer.rs:207); every (pixel, sample, bounce, draw slot) has its own stream.  All eight
 * RenderModes are rendered; ray_branching must be 1. */
enum {
    RAYNA_MODE_PBR = 0,             /* physically based path tracing */
    RAYNA_MODE_RAY_NORMAL = 1,      /* the normal against the ray, as (n / 2 + 0.5) */
    RAYNA_MODE_OUTWARD_NORMAL = 2,  /* the outward normal, as (n / 2 + 0.5) */
    RAYNA_MODE_SCATTER = 3,         /* the material's scatter direction, as (d / 2 + 0.5) */
    RAYNA_MODE_FRONT_FACE = 4,      /* white for back faces, red for front faces */
    RAYNA_MODE_DISTANCE = 5,        /* a palette ramp over 2 cbrt(distance) */
    RAYNA_MODE_UV = 6,              /* (u, v, 0), clamped */
    RAYNA_MODE_SIDE = 7             /* a palette colour per side index */
};  /* RenderMode (render_opts.rs:32-52), rendered as render_px_once (renderer.rs:410-490) */

typedef struct rayna_render_opts {
    uint32_t width;
    uint32_t height;
    uint32_t samples;
    uint32_t ray_depth;
    uint32_t ray_branching;
    int32_t mode;
    uint64_t seed;
} rayna_render_opts;

/* ---------------------------------------------------------------- stats (render/render.rs:4-15, extended) */
typedef struct rayna_render_stats {
    double duration_s;       /* host wall time of the render call */
    double kernel_ms;        /* sum of device time of all kernels of the frame (HIP events) */
    double extend_ms;        /* device time of the extend (closest-hit) kernels */
    double shade_ms;         /* device time of the shade kernels */
    uint64_t accum_frames;   /* frames in the accumulation buffer, as AccumulationBuffer::frame_count */
    uint64_t samples;        /* camera samples traced this frame (width * height * samples) */
    uint64_t segments;       /* closest-hit queries this frame (one per path vertex) */
    uint64_t extend_launches;
    double extend_camera_ms; /* of extend_ms: the launches for the camera rays (depth 0) */
} rayna_render_stats;

/* Output layouts of rayna_render. */
enum {
    RAYNA_LAYOUT_XMAJOR = 0,  /* [x][y][rgb], the engine's Image shape (w, h) (core/image.rs:24-26) */
    RAYNA_LAYOUT_ROWMAJOR = 1 /* [y][x][rgb] */
};

typedef struct rayna_renderer rayna_renderer;

/* Renderer::new_from (render/renderer.rs:86-107).  Builds the SAH BVHs on the host with the engine's
 * algorithm (shared/generic_bvh.rs:45-290), uploads the scene and allocates the path queues. */
rayna_status rayna_create(const rayna_scene_desc *scene, const rayna_camera *camera,
                          const rayna_render_opts *opts, rayna_renderer **out);

/* Renderer::set_scene / set_camera / set_options (render/renderer.rs:158-176): each clears accumulation. */
rayna_status rayna_set_scene(rayna_renderer *r, const rayna_scene_desc *scene);
rayna_status rayna_set_camera(rayna_renderer *r, const rayna_camera *camera);
rayna_status rayna_set_options(rayna_renderer *r, const rayna_render_opts *opts);

/* Renderer::clear_accumulation (render/renderer.rs:153). */
rayna_status rayna_clear_accumulation(rayna_renderer *r);

/* Renderer::render (render/renderer.rs:224-264): one frame of `samples` samples per pixel, accumulated
 * into the running per-pixel mean (render/accum_buffer.rs:31-45).  out_rgb receives width*height*3
 * floats in `layout`, or may be NULL to leave the image on the device (benchmarks).  An invalid camera
 * yields the engine's magenta/black checkerboard of ilog2(w) x ilog2(h) pixels (renderer.rs:266-292):
 * then `out_dims` (if not NULL) receives that size and out_rgb must hold width*height*3 floats. */
rayna_status rayna_render(rayna_renderer *r, float *out_rgb, int32_t layout, uint32_t out_dims[2],
                          rayna_render_stats *stats);

/* rayna_render followed by the UI's conversion of the image to 8 bits (rayna_ui/src/ext/img_ext.rs:15-45):
 * every channel raised to 1/2.2, scaled by 255 and cast to u8 with Rust's saturating cast, as RGBA (alpha
 * 255) in row-major order; out_rgba holds width*height*4 bytes (the checkerboard's size for an invalid
 * camera, as rayna_render).  The conversion runs on the GPU. */
rayna_status rayna_render_srgb8(rayna_renderer *r, uint8_t *out_rgba, uint32_t out_dims[2], rayna_render_stats *stats);

/* Replaces the colour of materials [0, n) of the current scene: the albedo of lambertian, metal,
 * dielectric and isotropic materials, the emission of lights (3 floats each, in rayna_scene_desc.materials
 * order).  Clears accumulation.  No engine counterpart: the engine fixes a material's texture when the scene
 * is built (material/lambertian.rs:12-14 etc.: the albedo is a texture held by the material), so a caller there rebuilds the Scene and calls set_scene
 * (render/renderer.rs:158-162); this keeps the BVH and the device scene and rewrites only the colours. */
rayna_status rayna_set_material_colours(rayna_renderer *r, const float *rgb, uint32_t n);

/* Path replay backpropagation (Vicini, Speierer and Jakob 2021) of frame `frame` since the last clear
 * (0 = the first; the frame's counter-RNG key, so the replay retraces exactly the paths that frame traced).
 * d_image is dloss/d(the frame's per-pixel sample mean), width*height*3 floats in `layout`; grad_rgb
 * receives dloss/d(material colour), n_materials*3 doubles (n_materials must be the scene's count).  Paths
 * are recomputed from the RNG, never stored; when the last rayna_render traced this same frame in one chunk
 * of paths and no setter ran since, its per-sample radiance is reused as the primal pass.  A channel whose
 * colour is exactly 0 receives no gradient through scattering.  No engine counterpart (the engine is not
 * differentiable); `stats` reports the call's device time and closest-hit queries. */
rayna_status rayna_adjoint(rayna_renderer *r, uint32_t frame, const float *d_image, int32_t layout,
                           double *grad_rgb, uint32_t n_materials, rayna_render_stats *stats);

/* ---------------------------------------------------------------- tuning */

/* Launch shape of the wavefront kernels.  Zero fields keep the default. */
typedef struct rayna_tuning {
    uint32_t chunk_slots;          /* paths in flight per chunk (default 2^27 for a renderer, 2^29 for a batch) */
    uint32_t refill_lanes;         /* extend: idle lanes of a wave before it takes new rays, 1..64 (default 32) */
    uint32_t wave_batch;           /* extend: queue positions a wave claims per atomic, 1..4096 (default 128) */
    uint32_t shade_blocks_per_cu;  /* shade: blocks of 256 threads per compute unit, 1..64 (default 16) */
    int32_t use_graph;             /* batches: capture the launch sequence in a hipGraph and replay it (1) */
    int32_t schedule;              /* batches: 0 = requests in the given order, 1 = fewest samples first,
                                    * 2 = grouped by scene (groups by total samples, ascending), fewest
                                    * samples first within a group: a wave then rarely mixes scenes,
                                    * 3 = interleaved: as 2, with chunks of one scene each, dealt in turn
                                    * from the scene groups so that concurrent pipelines hold different
                                    * scenes (eager only: with use_graph, RAYNA_UNSUPPORTED) */
    int32_t pipelines;             /* batches: chunks in flight at once, each on its own stream with its own
                                    * path state (chunk_slots paths each), 1..4 (default 3); the images do not
                                    * depend on it */
    int32_t packet_depths;         /* extend: bounces traced as wave packets when the chunk's samples of a pixel
                                    * are neighbouring slots (at least 16 per pixel, or a batch): 1 = the camera
                                    * rays only, d = the bounces before depth d, -1 = none, 0 = the default (1);
                                    * up to 16; the images do not depend on it */
} rayna_tuning;

/* Applies a tuning to a renderer.  A new chunk_slots re-plans the path buffers and clears accumulation. */
rayna_status rayna_set_tuning(rayna_renderer *r, const rayna_tuning *t);

/* rayna_adjoint without allocating: no counterpart in the engine; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  d_image_dev
 * (height x width x 3 floats, row-major) and grad_dev (3 doubles per material, overwritten) are device
 * memory on the renderer's device.  The gradient copies live in a static workspace of 4 MiB per GPU (the
 * gradients of at most 174762 materials, else RAYNA_INVALID_ARGUMENT), held until the process ends; the
 * renderer's own buffers must be in place (a frame rendered since the last set_scene / set_options /
 * set_tuning, else RAYNA_INVALID_ARGUMENT).  Calls on one GPU share the workspace: calls that overlap on one
 * GPU race.  Synchronous; no stats. */
rayna_status rayna_adjoint_static(rayna_renderer *r, uint32_t frame, const float *d_image_dev, double *grad_dev,
                                  uint32_t n_materials);

/* ---------------------------------------------------------------- debug validation (shared/validate.rs:13-198)
 * The engine asserts these invariants in debug builds along every path (render/renderer.rs:395-561).  With
 * validation on, the PBR frames of rayna_render count, per check, the rays, hits and colours that break them,
 * with the engine's tolerances (a unit vector: |n.n - 1| <= 2e-4, glam's is_normalized; a hit on its ray:
 * within 1e-6 absolute or 1e-3 relative, approx's relative_eq with validate.rs's EPSILON and RELATIVE); the
 * kernels never trap.  The counts equal those of the oracle's fp32 restatement bit for bit. */
enum {
    RAYNA_CHECK_RAY = 0,         /* a traced ray: origin not finite, or direction not finite and unit */
    RAYNA_CHECK_HIT_DIST = 1,    /* a hit's distance not finite or outside [0.001, FLT_MAX] */
    RAYNA_CHECK_HIT_POS = 2,     /* a hit's position not finite, or not the ray's point at its distance */
    RAYNA_CHECK_HIT_NORMAL = 3,  /* a hit's normal or ray normal not finite and unit */
    RAYNA_CHECK_HIT_UV = 4,      /* a hit's uv NaN or outside [0, 1] */
    RAYNA_CHECK_COLOUR = 5,      /* a material colour, emitted radiance (lights, sky) or path throughput NaN,
                                  * infinite or negative */
    RAYNA_CHECK_SAMPLE = 6,      /* a sample's radiance NaN, infinite or negative */
    RAYNA_CHECK_PIXEL = 7,       /* a pixel's sample mean NaN, infinite or negative */
    RAYNA_CHECKS = 8
};

/* Turns validation on (1) or off (0) for the renderer's later PBR frames (not batches, adjoints or the
 * non-PBR modes); turning it on clears the counts.  Validation adds a kernel variant and per-hit work. */
rayna_status rayna_set_validate(rayna_renderer *r, int32_t enable);

/* The counts of each RAYNA_CHECK_* since validation was turned on or last reset (counts: RAYNA_CHECKS
 * entries); reset != 0 clears them after reading. */
rayna_status rayna_validate_counts(rayna_renderer *r, uint64_t *counts, int32_t reset);

/* ---------------------------------------------------------------- batches
 * Many render requests (each its own scene, camera, resolution, samples and seed) traced together: the
 * requests' samples are packed into chunks of paths, and every chunk runs one generate / extend / shade
 * pipeline over the paths of all its requests, so small requests share launches and fill the GPU.  Each
 * request's image is bit for bit the image of a fresh Renderer's first frame with the same inputs.  No engine
 * counterpart (the engine renders one scene per Renderer::render call, render/renderer.rs:224-264). */
typedef struct rayna_batch rayna_batch;

typedef struct rayna_request {
    int32_t scene;            /* index into the scenes given to rayna_batch_create */
    int32_t layout;           /* RAYNA_LAYOUT_* of out_rgb */
    rayna_camera camera;
    rayna_render_opts opts;   /* mode RAYNA_MODE_PBR, ray_branching 1 */
    float *out_rgb;           /* host buffer of width*height*3 floats, or NULL to leave the image on the device
                               * (an invalid camera gives the engine's checkerboard, as rayna_render) */
} rayna_request;

typedef struct rayna_batch_stats {
    double duration_s;        /* host wall time of the call */
    double kernel_ms;         /* device time of the batch */
    uint64_t samples;         /* camera samples traced */
    uint64_t segments;        /* closest-hit queries */
    uint32_t chunks;          /* chunks of paths the batch was cut into */
    uint32_t launches;        /* kernel launches */
    double *request_ms;       /* if not NULL (n entries): device time from the batch start until each request's
                               * image was final; filled when the batch ran without a graph */
} rayna_batch_stats;

/* Packs and uploads the scenes (each as rayna_create does: the engine's SAH trees, flattened). */
rayna_status rayna_batch_create(const rayna_scene_desc *scenes, int32_t n_scenes, rayna_batch **out);
rayna_status rayna_batch_set_tuning(rayna_batch *b, const rayna_tuning *t);
/* Renders n requests; returns when every image is final (and copied to its out_rgb). */
rayna_status rayna_render_batch(rayna_batch *b, const rayna_request *reqs, int32_t n, rayna_batch_stats *stats);
void rayna_batch_destroy(rayna_batch *b);
const char *rayna_batch_last_error(const rayna_batch *b);

/* Parity hook (no engine counterpart; the engine's RenderMode debug views are the closest,
 * renderer.rs:428-490): closest hit of the camera ray of sample `sample_index` of every pixel, row-major.
 * prim_rank receives the primitive's position in the depth-first order of the engine's BVH
 * (-1 on a miss, -2 - i for the i-th unbounded object), t the hit distance. */
rayna_status rayna_primary_hits(rayna_renderer *r, uint32_t sample_index, int32_t *prim_rank, float *t);

/* Position of each bounded primitive in depth-first order: object index and index within the object's
 * mesh (triangle index within a TRIANGLE_LIST, else 0).  Returns the count via n_out. */
rayna_status rayna_dfs_order(rayna_renderer *r, int32_t *object, int32_t *local, int32_t capacity,
                             int32_t *n_out);

/* FNV-1a hash of the engine BVH topology the renderer built (TLAS and every BLAS). */
rayna_status rayna_tree_hash(rayna_renderer *r, uint64_t *hash);

/* Host-only scene preparation, as rayna_create performs it, without touching a GPU (no engine counterpart;
 * it exposes the BVH construction of shared/generic_bvh.rs:45-290 for inspection): the topology hash, the
 * depth-first order of bounded primitives (object, local) up to `capacity` entries, their count, and the
 * deepest traversal stack the flattened tree can need. */
rayna_status rayna_scene_prepare(const rayna_scene_desc *scene, uint64_t *tree_hash, int32_t *object,
                                 int32_t *local, int32_t capacity, int32_t *n_ranks, int32_t *stack_need);

/* Host-only, as rayna_scene_prepare: the flattened traversal tree the extend kernels walk (no engine
 * counterpart; for inspection of its stack bounds).  node_words: 8 words per node (the four child words, then
 * the four info words: INFO_EMPTY = -1 unused, 0 inner node, else count | kind << 4 with kind 1 = triangles,
 * 2 = items); item_words: 2 per item (kind, index: kind 6 = a mesh-BVH link to node `index`, 8 = a transformed
 * object whose entry is in xform_words); xform_words: 2 per transformed object (inner kind, inner index).  Any
 * array may be NULL (then only counted); counts[3] = nodes, items, transformed objects.  root: the root node
 * (-1: none); stack_need: the deepest per-lane stack the per-lane traversal can need; packet_need: the deepest
 * wave stack of the camera-ray packets. */
rayna_status rayna_scene_tree(const rayna_scene_desc *scene, int32_t *node_words, int32_t node_cap,
                              int32_t *item_words, int32_t item_cap, int32_t *xform_words, int32_t xform_cap,
                              int32_t counts[3], int32_t *root, int32_t *stack_need, int32_t *packet_need);

/* Renderer drop. */
void rayna_destroy(rayna_renderer *r);

/* Measurement hook (no engine counterpart): the fp32 VALU issue rate the current GPU sustains, the peak the
 * traversal kernels' VALU roofline is taken against.  A calibration kernel runs eight independent multiply-add
 * chains per lane (one v_mul_f32 and one v_add_f32 per link: no packed or fused forms), eight waves on every
 * SIMD, `iters` links per chain (1..2^20).  *wave_insts_per_s: wave64 VALU instructions per second, counted by
 * construction over one timed launch after a warm-up; *seconds (may be NULL): that launch's duration.  Errors
 * are reported through rayna_last_error(NULL). */
rayna_status rayna_measure_valu_peak(uint32_t iters, double *wave_insts_per_s, double *seconds);

/* Verification hook (no engine counterpart): the number of float bit patterns in [first, first + count) on which
 * the SDF distances' square root (sqrtf's sequence without its final pass-through of zeros and infinities, in
 * both its per-lane and its wave-scaled form) differs from the correctly rounded sqrtf, NaNs comparing equal; 0
 * means they agree bit for bit there.  count may be up to 2^32 - first (a uint64_t so that the whole range fits
 * one call).  Errors are reported through rayna_last_error(NULL). */
rayna_status rayna_debug_sqrt_mismatches(uint32_t first, uint64_t count, uint64_t *mismatches);

/* Debug fault injection.  No counterpart in the engine, whose render cannot fail (render/renderer.rs:224-264);
 * it is for the callers' recovery tests.  countdown >= 0 makes the launch check that many checks from now (0: the
 * next one) of any renderer in the process report a launch failure: that call returns RAYNA_DEVICE_ERROR with
 * "(injected)" in rayna_last_error and leaves its renderer lost (RAYNA_RENDERER_LOST from then on).  The kernels
 * themselves run as usual and no GPU fault is made.  countdown < 0 switches injection off.  rayna_render,
 * rayna_adjoint, rayna_adjoint_static and rayna_primary_hits make one launch check per call (rayna_render none for
 * an invalid camera's checkerboard); rayna_render_srgb8 makes its rayna_render's and one of its own.
 * Makes no HIP call; returns the previous setting (-1 when off). */
int32_t rayna_debug_fail_launches(int32_t countdown);

/* Text of the last error on this handle (or of the last failed rayna_create when r is NULL). */
const char *rayna_last_error(const rayna_renderer *r);

#ifdef __cplusplus
}
#endif

#endif /* RAYNA_GPU_H */
