// Device-resident scene layout of the gfx950 renderer (shared by the host packer and the HIP kernels).
//
// One BVH4 holds the engine's whole two-level hierarchy: the object BVH (object/bvh.rs) whose leaves are
// objects, and per triangle-list object its mesh BVH (mesh/advanced/bvh.rs).  The topology is the engine's
// SAH tree (shared/generic_bvh.rs:97-154: inner nodes have 2-4 children, leaf clusters up to 8 items); the
// boxes are refitted tight and rounded outward to float.  Every node is 128 bytes (one L2 line on CDNA):
// four child boxes as SoA float4s plus four child words.
#pragma once

#include <stdint.h>

#ifndef RAYNA_HD
#define RAYNA_HD
#endif

namespace rayna {

// Child word encoding (GNode::child / GNode::info):
//   info == INFO_EMPTY                  unused slot (its box is empty)
//   info == INFO_INNER                  child is GNode `child`
//   (info & 0xf) = count (1..8), (info >> 4) = LEAF_TRIS or LEAF_ITEMS: a leaf range of `count` entries
//   starting at `child` in the triangle array (LEAF_TRIS) or the item array (LEAF_ITEMS).
enum : int32_t { INFO_EMPTY = -1, INFO_INNER = 0 };
enum : int32_t { LEAF_TRIS = 1, LEAF_ITEMS = 2 };

// Traversal stack entry codes of the extend kernel: >= 0 an inner node; < 0 a leaf range (bit 30: item range,
// bits 27-29: count - 1, bits 0-26: first triangle or item); ENTRY_NONE marks an unused child slot.  The device
// copy of the nodes holds each child slot's code in `info` (see encode_child_codes), so an expansion reads
// boxes and codes only.
constexpr uint32_t LEAF_CODE_FLAG = 0x80000000u;
constexpr uint32_t LEAF_CODE_ITEMS = 0x40000000u;
constexpr uint32_t LEAF_CODE_START = 0x07ffffffu;
constexpr int32_t ENTRY_NONE = 0x7fffffff;
constexpr int32_t leaf_code(int32_t start, int count, bool items) {
    return (int32_t)(LEAF_CODE_FLAG | (items ? LEAF_CODE_ITEMS : 0u) | ((uint32_t)(count - 1) << 27) |
                     ((uint32_t)start & LEAF_CODE_START));
}
constexpr int32_t child_code(int32_t child, int32_t info) {
    return info == INFO_EMPTY   ? ENTRY_NONE
           : info == INFO_INNER ? child
                                : leaf_code(child, info & 0xf, (info >> 4) == LEAF_ITEMS);
}

struct alignas(16) GNode {
    float lo_x[4], lo_y[4], lo_z[4];
    float hi_x[4], hi_y[4], hi_z[4];
    int32_t child[4];
    int32_t info[4];
};
static_assert(sizeof(GNode) == 128, "GNode must be 128 bytes");

// Triangle record, 48 bytes: v0 + rank, e1 = v1 - v0 + material, e2 = v2 - v0 + flags, all float32.
// `flags` bit 0: the interpolated vertex normal may fail to normalise somewhere on the triangle, so a hit must
// check it (Triangle::intersect returns None then, triangle.rs:97).
struct alignas(16) GTri {
    float v0x, v0y, v0z;
    int32_t rank;
    float e1x, e1y, e1z;
    int32_t material;
    float e2x, e2y, e2z;
    int32_t flags;
};
static_assert(sizeof(GTri) == 48, "GTri must be 48 bytes");

// Vertex normals of a triangle (shading only), 48 bytes.
struct alignas(16) GTriNormals {
    float n0x, n0y, n0z, pad0;
    float n1x, n1y, n1z, pad1;
    float n2x, n2y, n2z, pad2;
};

// Item kinds (leaf entries of the object level and the kind field of a hit reference).
enum : int32_t {
    ITEM_NONE = 0,
    ITEM_TRIANGLE = 1,      // index: triangle array
    ITEM_SPHERE = 2,        // index: sphere array
    ITEM_PARALLELOGRAM = 3, // index: parallelogram array
    ITEM_AXIS_BOX = 4,      // index: box array
    ITEM_VOLUME = 5,        // index: volume array
    ITEM_LINK = 6,          // index: GNode (the mesh BVH of a triangle-list object)
    ITEM_SDF = 7,           // hit references only: index = unbounded SDF object
    ITEM_XFORM = 8          // index: transformed-object array (an object with a non-identity ObjectTransform)
};

struct alignas(16) GItem {
    int32_t kind;
    int32_t index;
    int32_t rank;
    int32_t material;
};

struct alignas(16) GSphere {
    float cx, cy, cz, radius;
};

// Planar constants computed in double on the host and rounded (planar/mod.rs:79-90).
struct alignas(16) GPara {
    float nx, ny, nz, d;
    float wx, wy, wz, pad0;
    float px, py, pz, pad1;
    float ux, uy, uz, pad2;
    float vx, vy, vz, pad3;
};

// AxisBoxMesh constants (axis_box.rs:29-37), computed in double and rounded.
struct alignas(16) GBox {
    float cx, cy, cz, pad0;
    float rx, ry, rz, pad1;
    float ix, iy, iz, pad2;
};

// A constant-density volume (object/volumetric.rs): boundary primitive + free-flight constant.
struct alignas(16) GVolume {
    int32_t boundary_kind;   // ITEM_SPHERE or ITEM_AXIS_BOX
    int32_t boundary_index;
    int32_t object;          // object index: selects the free-flight RNG stream
    float neg_inv_density;   // -1 / density, in float
};

// An unbounded SDF object (mesh/isosurface/raymarched.rs), enumerated kinds.
struct alignas(16) GSdf {
    int32_t kind;            // RAYNA_SDF_*
    int32_t first_sphere;    // into the sphere array
    int32_t n_spheres;
    int32_t max_iterations;
    int32_t bulb_iterations;
    int32_t power;
    int32_t material;
    int32_t list_pos;        // position among the scene's unbounded objects (the engine's tie order)
    float epsilon, k, bailout, pad1;
    float cx, cy, cz, pad2;
    // Escape sphere: a march point farther than escape_r from (ex, ey, ez) whose ray moves away from it can
    // never again see |distance| < epsilon (infinite: no such bound).
    float ex, ey, ez, escape_r;
};

// An object with a non-identity ObjectTransform (object/transform.rs): its primitive or volume in object space,
// and the affine maps both ways, row-major 3x4 ([m | t]: p -> m p + t), computed in double and rounded.
struct alignas(16) GXform {
    int32_t inner_kind;   // ITEM_SPHERE, ITEM_TRIANGLE, ITEM_PARALLELOGRAM, ITEM_AXIS_BOX or ITEM_VOLUME
    int32_t inner_index;  // into that kind's array
    int32_t pad0, pad1;
    float fwd[12];        // object -> world
    float inv[12];        // world -> object
};
static_assert(sizeof(GXform) == 112, "GXform must be 112 bytes");

// Material (material/*.rs), parameters rounded to float.  texture: -1 for the solid colour r, g, b, else the
// index of its texture.
struct alignas(16) GMaterial {
    int32_t kind;
    float r, g, b;
    float fuzz, ior, density;
    int32_t texture;
};

// A texture (texture/*.rs), parameters rounded to float.  odd / even / image are indices into the texture and
// image arrays; perm[i] is the offset of a 512-entry permutation table (noise).
struct alignas(16) GTexture {
    int32_t kind, odd, even, image;
    float r, g, b, scale;
    float ox, oy, oz, sx;
    float sy;
    int32_t perm[3];
    int32_t rgb, pad0, pad1, pad2;
};
static_assert(sizeof(GTexture) == 80, "GTexture must be 80 bytes");

// An image: width x height RGB texels (3 floats each, row-major) from `texel` on in the texel array.
struct alignas(16) GImage {
    int32_t texel;
    uint32_t width, height;
    int32_t pad;
};

// Hit record written by the extend kernel: distance, packed reference (kind << 28 | index), rank, aux bits.
enum : uint32_t { HIT_KIND_SHIFT = 28, HIT_INDEX_MASK = 0x0fffffffu };
// Hit flags of SDF hits: bit 0, the final distance sample was negative (front_face false); from bit 8, the march
// iteration of the hit (Intersection::side).  For an ITEM_XFORM hit the aux word holds the bits of the float
// distance along the object-space ray.
enum : uint32_t { HIT_AUX_BACK = 1u, HIT_AUX_SIDE_SHIFT = 8u };

// Camera constants in float, from Camera::calculate_viewport evaluated in double (camera.rs:114-167).
struct GViewport {
    float pos[3], pixel_center[3], vu[3], vv[3], du[3], dv[3];
};

// Everything a kernel needs to know about the scene.
struct SceneView {
    const GNode *nodes;
    const GTri *tris;
    const GTriNormals *tri_normals;
    const GItem *items;
    const GSphere *spheres;
    const GPara *paras;
    const GBox *boxes;
    const GVolume *volumes;
    const GSdf *sdfs;
    const GMaterial *materials;
    const GXform *xforms;
    const GTexture *textures;
    const GImage *images;
    const float *texels;
    const int32_t *perms;
    int32_t root;        // -1: no bounded object
    int32_t n_sdfs;
    int32_t skybox;
    int32_t stack_need;   // deepest per-lane stack k_extend can need (speculative traversal)
    int32_t packet_need;  // deepest wave stack k_extend_packet can need (depth-first, no speculation)
    int32_t instances;   // 1 if some object is a transformed mesh BVH (an instance)
    int32_t sphere_tests;  // 1 if traversal intersects spheres (sphere items, sphere-bounded media)
    int32_t media;       // 1 if the scene holds constant-density media (volume items)
    int32_t textured;    // 1 if some material has a texture or the sky is an image
    int32_t sky_image;   // RAYNA_SKY_HDRI: the sky's image
};

}  // namespace rayna
