"""BVH meshes of primitives (BvhMesh, rayna_engine/src/mesh/advanced/bvh.rs) and instanced mesh trees (a mesh BVH
under an ObjectTransform) in the oracle and the host packer.

Pinned without the oracle's own formulas: a BvhMesh is one mesh whose closest hit is the closest of its members, so
a scene holding the members as separate objects of the same material renders the same image bit for bit; a
rotated instance of a ball cloud hits where the balls with rotated centres do; the packer builds the oracle's
trees and depth-first order for the engine's RTTNW preset (a BvhMesh floor and an instanced ball cube).
"""
import numpy as np
import pytest

import oracle
from rayna_amd import presets
from rayna_amd import _abi as A
from rayna_amd.renderer import RayError, prepare_scene
from rayna_amd.scene import (AxisBoxMesh, BvhMesh, Camera, LambertianMaterial, ParallelogramMesh,
                             RaymarchedIsosurfaceMesh, RenderOpts, Scene, SdfSphere, SimpleObject, SimpleSkybox,
                             SphereMesh, Transform3, TriangleListMesh, TriangleMesh, VolumetricObject,
                             IsotropicMaterial, object_matrix)


def _members(seed=3, n=40):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        c = rng.uniform(-3, 3, size=3) * np.array([1.0, 0.4, 1.0]) + np.array([0.0, 1.0, 0.0])
        kind = k % 4
        if kind == 0:
            out.append(SphereMesh(tuple(c), float(rng.uniform(0.2, 0.5))))
        elif kind == 1:
            out.append(AxisBoxMesh(tuple(c - 0.3), tuple(c + rng.uniform(0.1, 0.5, size=3))))
        elif kind == 2:
            v = c + rng.normal(size=(3, 3)) * 0.4
            n = np.cross(v[1] - v[0], v[2] - v[0])
            n = n / np.linalg.norm(n)
            out.append(TriangleMesh(v, np.stack([n, n, n])))
        else:
            out.append(ParallelogramMesh(tuple(c), tuple(rng.normal(size=3) * 0.5), tuple(rng.normal(size=3) * 0.5)))
    return out


def _view():
    cam = Camera(pos=(0.0, 3.0, 9.0), fwd=(0.0, -0.3, -1.0), v_fov=50.0, focus_dist=8.0)
    return cam, RenderOpts(width=48, height=32, samples=3, ray_depth=5, seed=4)


def test_bvh_mesh_renders_as_its_members():
    members = _members()
    mat = LambertianMaterial((0.6, 0.5, 0.4))
    ground = SimpleObject(SphereMesh((0.0, -100.0, 0.0), 99.5), LambertianMaterial((0.5, 0.5, 0.5)))
    one = Scene([SimpleObject(BvhMesh(members), mat), ground], SimpleSkybox)
    many = Scene([SimpleObject(m, mat) for m in members] + [ground], SimpleSkybox)
    cam, opts = _view()
    for prec, box in ((32, oracle.BOX_TIGHT), (64, oracle.BOX_REFERENCE)):
        a, sa, _, _ = oracle.TwinScene(one).render(cam, opts, precision=prec, box_mode=box)
        b, sb, _, _ = oracle.TwinScene(many).render(cam, opts, precision=prec, box_mode=box)
        assert np.array_equal(a, b) and sa == sb
    # ranks: one per member, depth-first through the mesh's own tree, with the member index as `local`
    _, obj, loc, _ = prepare_scene(one)
    assert sorted(loc[obj == 0]) == list(range(len(members)))


def test_instanced_ball_cloud_is_the_rotated_cloud():
    rng = np.random.default_rng(8)
    centres = rng.random((200, 3)) * 0.825
    T = Transform3.from_axis_angle((0.0, 1.0, 0.0), 15.0).then_translate((-0.2, 0.5, 0.1))
    inst = SimpleObject(BvhMesh([SphereMesh(tuple(c), 0.1) for c in centres]), LambertianMaterial(), T,
                        corrected=False)
    M = T.matrix
    moved = SimpleObject(BvhMesh([SphereMesh(tuple(M[:3, :3] @ c + M[:3, 3]), 0.1) for c in centres]),
                         LambertianMaterial())
    ta, tb = oracle.TwinScene(Scene([inst], SimpleSkybox)), oracle.TwinScene(Scene([moved], SimpleSkybox))
    la, lb = ta.dfs_order()[1], tb.dfs_order()[1]  # rank -> ball index (the trees differ, the balls do not)
    hits = 0
    for _ in range(300):
        o = np.array([0.2, 0.9, -4.0]) + rng.normal(size=3) * 0.05
        d = rng.normal(size=3) * 0.12 + np.array([0.0, 0.0, 1.0])
        d /= np.linalg.norm(d)
        ha, hb = ta.trace(o, d), tb.trace(o, d)
        assert (ha is None) == (hb is None)
        if ha is None:
            continue
        hits += 1
        assert abs(ha["t"] - hb["t"]) < 1e-12 and la[ha["rank"]] == lb[hb["rank"]]
        assert np.allclose(ha["normal"], hb["normal"], atol=1e-12)
    assert hits > 100


def test_instanced_triangle_list_is_the_moved_list():
    tv, tn, _ = presets.icosphere(2, seed=1)
    T = Transform3.from_axis_angle(np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5]), 70.0)
    inst = SimpleObject(TriangleListMesh(tv, tn), LambertianMaterial(), T, corrected=False)
    R = T.matrix[:3, :3]
    moved = SimpleObject(TriangleListMesh(tv @ R.T, tn @ R.T), LambertianMaterial())
    ta, tb = oracle.TwinScene(Scene([inst], SimpleSkybox)), oracle.TwinScene(Scene([moved], SimpleSkybox))
    rng = np.random.default_rng(2)
    hits = 0
    for _ in range(200):
        o = np.array([0.0, 0.0, -4.0]) + rng.normal(size=3) * 0.1
        d = rng.normal(size=3) * 0.1 + np.array([0.0, 0.0, 1.0])
        d /= np.linalg.norm(d)
        ha, hb = ta.trace(o, d), tb.trace(o, d)
        if ha is None or hb is None:
            assert ha is None and hb is None
            continue
        hits += 1
        assert abs(ha["t"] - hb["t"]) < 1e-12
        assert np.allclose(ha["normal"], hb["normal"], atol=1e-10)
    assert hits > 100


def test_packer_builds_the_oracle_trees_for_rttnw():
    sc, _, _ = presets.rttnw(balls=300)
    h, obj, loc, need = prepare_scene(sc)
    ts = oracle.TwinScene(sc)
    assert h == ts.tree_hash()
    o2, l2 = ts.dfs_order()
    assert np.array_equal(obj, o2) and np.array_equal(loc, l2)
    assert len(obj) == 400 + 7 + 300  # every floor box, every ball and the seven single primitives
    assert need >= 6  # an instance runs four entries above its leaf


def test_bvh_mesh_centre_is_the_mean_of_its_members():
    sc, _, _ = presets.rttnw(balls=50)
    cube = sc.objects[-1]
    c = np.mean([np.asarray(m.pos) for m in cube.mesh.meshes], axis=0)
    M = object_matrix(cube)
    assert np.allclose(M[:3, :3] @ c + M[:3, 3], c + np.array([-1.0, 2.7, 3.95]) + 0.825, atol=1e-12)


def test_unsupported_bvh_meshes_are_refused():
    sdf = RaymarchedIsosurfaceMesh(SdfSphere((0, 0, 0), 1.0))
    ground = SimpleObject(SphereMesh((0, -100, 0), 99), LambertianMaterial())
    with pytest.raises(TypeError):
        Scene([SimpleObject(BvhMesh([SphereMesh((0, 0, 0), 1), sdf]), LambertianMaterial()), ground]).desc()
    vol = VolumetricObject(BvhMesh([SphereMesh((0, 0, 0), 1)]), IsotropicMaterial(), 1.0)
    with pytest.raises(RayError) as e:
        prepare_scene(Scene([vol, ground], SimpleSkybox))
    assert e.value.status == A.RAYNA_UNSUPPORTED
    # a BVH mesh naming an SDF mesh through the C description
    sc = Scene([SimpleObject(sdf, LambertianMaterial()), SimpleObject(BvhMesh([SphereMesh((0, 0, 0), 1)]),
                                                                      LambertianMaterial())], SimpleSkybox)
    d = sc.desc()
    assert d.meshes[2].kind == A.MESH_BVH
    d.meshes[2].index = 0  # now it holds the SDF mesh
    with pytest.raises(RayError) as e:
        prepare_scene(sc)
    assert e.value.status == A.RAYNA_UNSUPPORTED
    # an empty BVH mesh never hits
    empty = Scene([SimpleObject(BvhMesh([]), LambertianMaterial()), ground], SimpleSkybox)
    assert oracle.TwinScene(empty).trace([0, 5, 0], [0, 1, 0]) is None
    prepare_scene(empty)
