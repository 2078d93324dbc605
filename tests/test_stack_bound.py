"""The traversal stacks never outgrow what the packer sizes for them.

k_extend keeps one stack per lane (top in registers, then LDS, then a per-thread overflow column sized from
the packer's `stack_need`); k_extend_packet keeps one wave stack of PACKET_STACK entries used when the
packer's `packet_need` fits.  These tests replay both disciplines on the CPU over the flattened tree the
packer emits (rayna_scene_tree, no GPU): random culling, random child order, and -- for k_extend -- a
speculation that a lane holding a leaf range carries on for as long as the rest of its wave keeps searching
(any length, up to a second range or an empty stack), after which the held range's mesh-BVH links and
instance entries go on top; a held triangle range (which pushes nothing) may also be kept, part-tested, through
further speculation, as the primitive loop's early end leaves it (a part-tested range of items keeps its lane from
speculating until it is done, so its entries still go on top as one block, as modelled).  The peak must stay within the reported bound on scenes built to be adversarial
(a mesh-BVH leaf next to a wide subtree, instances, nested BVH meshes) and on the presets.
"""
import numpy as np
import pytest

from rayna_amd import (BvhMesh, LambertianMaterial, Scene, SimpleObject, SimpleSkybox, SphereMesh, Transform3,
                       TriangleListMesh, presets)
from rayna_amd.renderer import scene_tree

INFO_EMPTY, INFO_INNER = -1, 0
LEAF_ITEMS = 2
ITEM_LINK, ITEM_XFORM = 6, 8
PACKET_STACK = 64


class Tree:
    def __init__(self, scene):
        self.nodes, self.items, self.xforms, self.root, self.stack_need, self.packet_need = scene_tree(scene)

    def children(self, g):
        out = []
        for k in range(4):
            info = int(self.nodes[g, 4 + k])
            if info == INFO_EMPTY:
                continue
            if info == INFO_INNER:
                out.append(("node", int(self.nodes[g, k])))
            else:
                out.append(("range", int(self.nodes[g, k]), info & 0xF, info >> 4))
        return out


def lane_peak(tree, rng, p_hit, p_go_on):
    """Peak stack height of one lane of k_extend (its `sp`, the top entry in registers included)."""
    stack = [("node", tree.root)]
    peak, held = 1, None

    def push(e):
        nonlocal peak
        stack.append(e)
        peak = max(peak, len(stack))

    while True:
        stalled = False
        # inner loop: a lane without a range searches on; one that holds a range speculates while its wave does
        while stack and not stalled:
            if held is not None and rng.random() > p_go_on:
                break
            e = stack.pop()
            if e[0] == "marker":  # an instance's end: a lane holding a range puts it back and waits
                if held is not None:
                    push(e)
                    stalled = True
                else:
                    del stack[-3:]  # finish_instance pops the three saved words
                continue
            if e[0] == "word":
                raise AssertionError("a saved closest-hit word was popped as an entry")
            if rng.random() > p_hit:  # culled against the closest hit so far
                continue
            if e[0] == "range":
                if held is not None:
                    push(e)
                    stalled = True
                else:
                    held = e
                continue
            kids = [c for c in tree.children(e[1]) if rng.random() < p_hit]
            rng.shuffle(kids)
            for c in kids:
                push(c)
        if held is None:
            if not stack:
                return peak
            continue
        _, start, count, kind = held
        if kind != LEAF_ITEMS:
            # a triangle range may be left part-tested when the primitive loop ends early; the lane keeps it
            # and speculates again
            if rng.random() < 0.5:
                continue
            held = None
            continue
        held = None
        for i in range(count):
            ik, idx = (int(v) for v in tree.items[start + i])
            if ik == ITEM_LINK:
                push(("node", idx))
            elif ik == ITEM_XFORM and int(tree.xforms[idx, 0]) == ITEM_LINK:
                if i + 1 < count:
                    push(("range", start + i + 1, count - i - 1, kind))
                for _ in range(3):
                    push(("word",))
                push(("marker",))
                push(("node", int(tree.xforms[idx, 1])))
                break


def packet_peak(tree, rng, p_hit):
    """Peak of k_extend_packet's wave stack: depth-first, children pushed when some lane hits them, mesh-BVH
    links of a leaf range pushed as the range is processed."""
    stack = [("node", tree.root)]
    peak = 1
    while stack:
        e = stack.pop()
        if e[0] == "node":
            kids = [c for c in tree.children(e[1]) if rng.random() < p_hit]
            rng.shuffle(kids)
            stack.extend(kids)
        else:
            _, start, count, kind = e
            if kind == LEAF_ITEMS:
                stack.extend(("node", int(tree.items[start + i, 1])) for i in range(count)
                             if int(tree.items[start + i, 0]) == ITEM_LINK)
        peak = max(peak, len(stack))
    return peak


def _mesh(rng, centre, n=40):
    v = rng.normal(size=(n, 3, 3)) * 0.3 + np.asarray(centre)
    nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    return TriangleListMesh(v, np.repeat(nrm[:, None, :], 3, axis=1))


def _adversarial(kind, seed=0):
    """A mesh-BVH leaf beside a wide subtree of spheres (the case where a held range's links land on top of a
    speculative descent), optionally with instances of the mesh and with BVH meshes of spheres."""
    rng = np.random.default_rng(seed)
    m = LambertianMaterial()
    objs = [SimpleObject(_mesh(rng, (0.0, 0.0, 0.0)), m)]
    for i in range(24):
        objs.append(SimpleObject(SphereMesh(tuple(rng.uniform(2.0, 9.0, 3)), 0.2), m))
    if kind in ("instances", "all"):
        for i in range(6):
            xf = Transform3.from_axis_angle((0.0, 1.0, 0.0), 30.0 * i).then_translate(tuple(rng.uniform(-8, -2, 3)))
            objs.append(SimpleObject(_mesh(rng, (0.0, 0.0, 0.0), 30), m, xf))
    if kind in ("bvh_meshes", "all"):
        for i in range(4):
            balls = [SphereMesh(tuple(rng.uniform(-3, 3, 3) + [0, 10, 0]), 0.1) for _ in range(20)]
            objs.append(SimpleObject(BvhMesh(balls), m))
    for i in range(8):
        objs.append(SimpleObject(_mesh(rng, tuple(rng.uniform(-9, 9, 3)), 12), m))
    return Scene(objs, SimpleSkybox)


SCENES = {
    "adversarial": lambda: _adversarial("plain"),
    "adversarial_instances": lambda: _adversarial("instances", 1),
    "adversarial_bvh_meshes": lambda: _adversarial("bvh_meshes", 2),
    "adversarial_all": lambda: _adversarial("all", 3),
    "mixed_primitives": lambda: presets.mixed_primitives(width=8, height=8, samples=1)[0],
    "rttnw": lambda: presets.rttnw(width=8, height=8, samples=1)[0],
    "cornell": lambda: presets.cornell(width=8, height=8, samples=1)[0],
    "icosphere_mesh": lambda: presets.triangle_mesh(subdiv=3, width=8, height=8, samples=1)[0],
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_lane_stack_stays_within_stack_need(name):
    tree = Tree(SCENES[name]())
    rng = np.random.default_rng(7)
    peak = 0
    for p_hit in (1.0, 0.9, 0.6):
        for p_go_on in (1.0, 0.97, 0.5, 0.0):
            for _ in range(60):
                peak = max(peak, lane_peak(tree, rng, p_hit, p_go_on))
    assert peak <= tree.stack_need, (peak, tree.stack_need)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_packet_stack_stays_within_packet_need(name):
    tree = Tree(SCENES[name]())
    rng = np.random.default_rng(8)
    peak = max(packet_peak(tree, rng, p) for p in (1.0, 0.9, 0.6) for _ in range(60))
    assert peak <= tree.packet_need <= tree.stack_need, (peak, tree.packet_need, tree.stack_need)


def test_speculation_can_exceed_the_depth_first_bound():
    """The case the speculative bound exists for: with every box hit and speculation carried to its limit, a
    lane's stack on the adversarial scene rises above the depth-first need (so a bound without the held range's
    entries would let the overflow column be written past its end)."""
    tree = Tree(SCENES["adversarial"]())
    rng = np.random.default_rng(9)
    peak = max(lane_peak(tree, rng, 1.0, 1.0) for _ in range(200))
    assert tree.packet_need < peak <= tree.stack_need, (tree.packet_need, peak, tree.stack_need)
