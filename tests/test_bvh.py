"""Pins the engine's SAH BVH construction (rayna_engine/src/shared/generic_bvh.rs:45-290) three ways:

* an independent Python restatement of the literal algorithm (every candidate split re-unions its slices,
  unions folded from Aabb::default() so they include the origin, ties in the axis sort by input index) against
  the oracle's C++ literal and sweep builders, on inputs with many equal sort keys;
* the product's presorted builder (rayna_scene_prepare, no GPU needed) against the oracle's trees: topology hash
  and the depth-first order of every primitive, up to the full 1.31M-triangle workload;
* structural invariants: every item once, nested nodes with 2..8 children, split nodes with at most 4.
"""
import numpy as np
import pytest

import oracle
from rayna_amd import (AxisBoxMesh, LambertianMaterial, Scene, SimpleObject, SimpleSkybox, TriangleListMesh,
                       prepare_scene, presets)


def _area(mn, mx):
    s = mx - mn
    return ((s[0] * s[1]) + (s[1] * s[2]) + (s[2] * s[0])) * 2.0


def _union(boxes, ids):
    mn, mx = np.zeros(3), np.zeros(3)  # Aabb::default()
    for i in ids:
        mn = np.minimum(mn, boxes[i][0])
        mx = np.maximum(mx, boxes[i][1])
    return mn, mx


def engine_tree(boxes):
    """generate_nodes_sah, restated literally.  Returns ('item', i) / ('nested', [children])."""
    def sort_axis(ids, axis):
        return sorted(ids, key=lambda i: (boxes[i][0][axis], i))

    def optimal_split(ids):
        n = len(ids)
        if n <= 2:
            return None
        best = None
        for axis in range(3):
            ids = sort_axis(ids, axis)
            for p in range(1, n - 1):
                cost = p * _area(*_union(boxes, ids[:p])) + (n - p) * _area(*_union(boxes, ids[p:]))
                if best is None or cost < best[2]:
                    best = (axis, p, cost)
        return best

    def generate(ids):
        if len(ids) == 1:
            return ("item", ids[0])
        if len(ids) <= 8:
            return ("nested", [("item", i) for i in ids])
        axis, p, _ = optimal_split(ids)
        ids = sort_axis(ids, axis)
        children = []
        for half in (ids[:p], ids[p:]):
            s = optimal_split(half)
            if s:
                half = sort_axis(half, s[0])
                children += [generate(half[: s[1]]), generate(half[s[1]:])]
            else:
                children.append(generate(half))
        return ("nested", children)

    return generate(list(range(len(boxes))))


def walk(tree):
    items, nch = [], []

    def rec(t):
        if t[0] == "item":
            items.append(t[1])
            nch.append(0)
        else:
            items.append(-1)
            nch.append(len(t[1]))
            for c in t[1]:
                rec(c)

    rec(tree)
    return np.array(items), np.array(nch)


def _box_scene(boxes):
    objs = [SimpleObject(AxisBoxMesh(tuple(b[0]), tuple(b[1])), LambertianMaterial()) for b in boxes]
    return Scene(objs, SimpleSkybox)


def _random_boxes(n, seed, offset=0.0):
    rng = np.random.default_rng(seed)
    lo = rng.integers(-6, 6, (n, 3)).astype(float) + offset  # integer corners: many equal sort keys
    size = rng.integers(1, 4, (n, 3)).astype(float)
    return [(lo[i], lo[i] + size[i]) for i in range(n)]


@pytest.mark.parametrize("n", [1, 2, 3, 8, 9, 17, 40, 150])
@pytest.mark.parametrize("offset", [0.0, 25.0])
def test_object_bvh_matches_literal_python(n, offset):
    boxes = _random_boxes(n, n * 7 + int(offset), offset)
    sc = _box_scene(boxes)
    want = walk(engine_tree(boxes))
    for builder in ("literal", "sweep"):
        items, nch = oracle.TwinScene(sc, builder).tree_walk(-1)
        assert np.array_equal(items, want[0]) and np.array_equal(nch, want[1]), builder


def test_mesh_bvh_matches_literal_python():
    rng = np.random.default_rng(11)
    tv = np.round(rng.uniform(-3, 3, (60, 3, 3)), 1) + [4.0, 0.0, 0.0]
    tn = np.tile([0.0, 0.0, 1.0], (60, 3, 1))
    sc = Scene([SimpleObject(TriangleListMesh(tv, tn), LambertianMaterial())], SimpleSkybox)
    boxes = [(t.min(axis=0), t.max(axis=0)) for t in tv]
    want = walk(engine_tree(boxes))
    items, nch = oracle.TwinScene(sc).tree_walk(0)
    assert np.array_equal(items, want[0]) and np.array_equal(nch, want[1])


def test_origin_is_part_of_every_union():
    """The origin folded into Aabb::encompass_iter changes the split the SAH picks; a builder that leaves it
    out builds a different tree for these boxes."""
    lo = [[9, 4, 6], [13, 0, 1], [12, 7, 1], [8, 6, 3], [8, 6, 2], [9, 5, 4], [6, 0, 6], [12, 6, 4], [12, 2, 3],
          [12, 0, 2]]
    boxes = [(np.array(c, float), np.array(c, float) + 1.0) for c in lo]
    with_origin = walk(engine_tree(boxes))
    global _union
    saved = _union

    def _union_no_origin(bx, ids):
        mn = np.min([bx[i][0] for i in ids], axis=0)
        mx = np.max([bx[i][1] for i in ids], axis=0)
        return mn, mx

    _union = _union_no_origin
    try:
        without = walk(engine_tree(boxes))
    finally:
        _union = saved
    assert not (np.array_equal(with_origin[0], without[0]) and np.array_equal(with_origin[1], without[1]))
    items, nch = oracle.TwinScene(_box_scene(boxes)).tree_walk(-1)
    assert np.array_equal(items, with_origin[0]) and np.array_equal(nch, with_origin[1])


def _invariants(items, nch, n_items):
    assert sorted(items[items >= 0].tolist()) == list(range(n_items))
    # rebuild parent/child structure from the preorder walk
    pos = 0

    def rec():
        nonlocal pos
        i = pos
        pos += 1
        if items[i] >= 0:
            return True
        kids = [rec() for _ in range(nch[i])]
        assert 2 <= nch[i] <= 8
        if not all(kids):  # has a nested child: a split node
            assert nch[i] <= 4
        return False

    rec()
    assert pos == len(items)


def test_tree_invariants_on_workloads():
    sc, _, _ = presets.random_spheres()
    items, nch = oracle.TwinScene(sc).tree_walk(-1)
    _invariants(items, nch, len(sc.objects))
    sc, _, _ = presets.triangle_mesh(subdiv=4)
    tw = oracle.TwinScene(sc)
    for m in range(21):
        items, nch = tw.tree_walk(m)
        _invariants(items, nch, len(sc.objects[m].mesh.vertices))


@pytest.mark.parametrize("make", [presets.random_spheres, presets.mixed_primitives, presets.sdf_volume,
                                  lambda: presets.triangle_mesh(subdiv=5)])
def test_product_builder_matches_oracle(make):
    sc, _, _ = make()
    h, obj, loc, need = prepare_scene(sc)
    tw = oracle.TwinScene(sc)
    o2, l2 = tw.dfs_order()
    assert h == tw.tree_hash()
    assert np.array_equal(obj, o2) and np.array_equal(loc, l2)
    assert need >= 1


def test_product_builder_matches_oracle_full_triangle_mesh():
    """The 1080p workload's 20 x 65,536-triangle mesh BVHs and its object BVH."""
    sc, _, _ = presets.triangle_mesh()
    h, obj, loc, _ = prepare_scene(sc)
    tw = oracle.TwinScene(sc)
    o2, l2 = tw.dfs_order()
    assert len(obj) == 1310720 + 2 + 8
    assert h == tw.tree_hash()
    assert np.array_equal(obj, o2) and np.array_equal(loc, l2)


def test_prepare_rejects_bad_scenes():
    from rayna_amd import RayError, SphereMesh
    from rayna_amd import _abi as A
    sc = Scene([SimpleObject(SphereMesh((0, 0, 0), float("nan")), LambertianMaterial())], SimpleSkybox)
    with pytest.raises(RayError) as e:
        prepare_scene(sc)
    assert e.value.status == A.RAYNA_INVALID_ARGUMENT
