// Host-side preparation of a scene for the GPU: the engine's object graph and SAH trees, flattened into the
// device layout of layout.hpp.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rayna_gpu.h"
#include "layout.hpp"

namespace rayna {

// One tree of shared/generic_bvh.rs: nodes are either an item (GenericBvhNode::Object) or nested, with children
// in insertion order.
struct RefTree {
    struct Node {
        int32_t item;          // >= 0: item index; -1: nested
        int32_t first_child;   // into `children`
        int32_t n_children;
    };
    std::vector<Node> nodes;
    std::vector<int32_t> children;
    int32_t root = -1;
};

struct Aabb64 {
    double mn[3], mx[3];
};

// The engine's SAH builder (generic_bvh.rs:45-290) over items with the given boxes, ties in the axis sorts
// broken by item index.
RefTree build_engine_bvh(const std::vector<Aabb64> &boxes);

// FNV-1a hash of a tree's depth-first walk (item or -1, child count per node), chained onto `h`.
void hash_tree(const RefTree &t, uint64_t &h);

struct PackedScene {
    std::vector<GNode> nodes;
    std::vector<GTri> tris;
    std::vector<GTriNormals> tri_normals;
    std::vector<GItem> items;
    std::vector<GSphere> spheres;
    std::vector<GPara> paras;
    std::vector<GBox> boxes;
    std::vector<GVolume> volumes;
    std::vector<GSdf> sdfs;
    std::vector<GMaterial> materials;
    std::vector<GXform> xforms;
    std::vector<GTexture> textures;
    std::vector<GImage> images;
    std::vector<float> texels;
    std::vector<int32_t> perms;
    int32_t root = -1;
    int32_t stack_need = 0;   // deepest per-lane stack of k_extend (speculative traversal)
    int32_t packet_need = 0;  // deepest wave stack of k_extend_packet (depth-first, no speculation)
    int32_t skybox = RAYNA_SKY_NONE;
    int32_t sky_image = 0;
    uint64_t tree_hash = 0;
    std::vector<int32_t> rank_object, rank_local;  // depth-first rank -> (object, index in its mesh)
};

// True if some object of the packed scene is an instance (a mesh BVH under a non-identity transform).
inline bool has_instances(const PackedScene &p) {
    for (const GXform &x : p.xforms)
        if (x.inner_kind == ITEM_LINK) return true;
    return false;
}

// True if traversal of the packed scene intersects spheres: a sphere item (also inside a BVH mesh), a medium
// bounded by a sphere, or a transformed sphere or sphere-bounded medium.
inline bool has_spheres(const PackedScene &p) {
    for (const GItem &it : p.items)
        if (it.kind == ITEM_SPHERE) return true;
    for (const GVolume &v : p.volumes)
        if (v.boundary_kind == ITEM_SPHERE) return true;
    for (const GXform &x : p.xforms)
        if (x.inner_kind == ITEM_SPHERE) return true;
    return false;
}

// True if some material of the packed scene has a texture or its sky is an image.
inline bool has_textures(const PackedScene &p) {
    if (p.skybox == RAYNA_SKY_HDRI) return true;
    for (const GMaterial &m : p.materials)
        if (m.texture >= 0) return true;
    return false;
}

// The permutation of an improved Perlin noise for `seed`: 0..255 shuffled by Fisher-Yates with SplitMix64 draws,
// repeated to 512 entries.
void noise_permutation(uint32_t seed, int32_t out[512]);

// Validates and packs a scene.  Returns RAYNA_OK, or an error status with a message in `error`.
rayna_status pack_scene(const rayna_scene_desc &d, PackedScene &out, std::string &error);

// Where one scene of a merged scene set lives.
struct SceneSlot {
    int32_t root;       // its tree's root node, -1 if it has no bounded object
    int32_t first_sdf;  // its unbounded objects [first_sdf, first_sdf + n_sdfs)
    int32_t n_sdfs;
    int32_t skybox;
    int32_t stack_need;
    int32_t packet_need;
    int32_t sky_image;
};

// Concatenates packed scenes into one device layout: every index (child nodes, leaf ranges, item targets,
// volume boundaries, SDF spheres, materials) is offset into the merged arrays.  Ranks stay per scene (they
// only order hits within one traversal).  Returns false if the merged arrays overflow a leaf or hit index.
bool merge_scenes(const std::vector<PackedScene> &in, PackedScene &out, std::vector<SceneSlot> &slots);

// The device copy of a node array: each child slot's `info` replaced by its traversal stack entry code
// (child_code: the inner node, the leaf range's code, or ENTRY_NONE).
std::vector<GNode> device_nodes(const std::vector<GNode> &nodes);

}  // namespace rayna
