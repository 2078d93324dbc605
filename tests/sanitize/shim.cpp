// tests/sanitize/shim.cpp -- C entry of the host scene packer (rayna_amd/csrc/scene_build.cpp) for the sanitizer
// builds: the packer and the oracle compiled with AddressSanitizer + UndefinedBehaviorSanitizer (or
// ThreadSanitizer) into one host-only library, driven from Python by tests/test_sanitized_host.py.
#include <string>

#include "../../rayna_amd/csrc/scene_build.hpp"

extern "C" int32_t san_pack_scene(const rayna_scene_desc *d, uint64_t *tree_hash, int32_t *stack_need,
                                  int32_t *packet_need, int32_t *n_nodes) {
    rayna::PackedScene p;
    std::string err;
    const rayna_status st = rayna::pack_scene(*d, p, err);
    if (st != RAYNA_OK) return st;
    *tree_hash = p.tree_hash;
    *stack_need = p.stack_need;
    *packet_need = p.packet_need;
    *n_nodes = (int32_t)p.nodes.size();
    // the device copy of the nodes and a merge of the scene with itself, as batches do
    const std::vector<rayna::GNode> dev = rayna::device_nodes(p.nodes);
    std::vector<rayna::PackedScene> two(2, p);
    rayna::PackedScene merged;
    std::vector<rayna::SceneSlot> slots;
    if (!rayna::merge_scenes(two, merged, slots)) return -1;
    return (int32_t)(dev.size() == p.nodes.size() && merged.nodes.size() == 2 * p.nodes.size() ? 0 : -2);
}
