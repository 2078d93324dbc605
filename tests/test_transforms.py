"""Object transforms (ObjectTransform, rayna_engine/src/object/transform.rs) in the oracle and the host packer.

The oracle's transformed intersections are pinned by closed forms computed here independently of it: the ray is
taken to object space with numpy's general matrix inverse, the primitive is intersected by the slab method or the
ray-quadric quadratic, and the record is mapped out the engine's way (the position by the map, the normal by the
map's linear part and normalised -- not by the inverse transpose -- and the distance measured from the world
ray's origin).  Host-only checks cover the packer: transformed boxes feed the same SAH trees as the oracle's,
identity matrices are the identity transform, and unsupported or invalid transforms are refused.
"""
import math

import numpy as np
import pytest

import oracle
from rayna_amd import presets
from rayna_amd import _abi as A
from rayna_amd.renderer import RayError, prepare_scene
from rayna_amd.scene import (AxisBoxMesh, LambertianMaterial, ParallelogramMesh, RaymarchedIsosurfaceMesh, Scene,
                             SdfSphere, SimpleObject, SimpleSkybox, SphereMesh, Transform3, TriangleListMesh,
                             TriangleMesh, VolumetricObject, IsotropicMaterial, object_matrix)

Y = (0.0, 1.0, 0.0)


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------------------------------ Transform3


def test_axis_angle_is_a_proper_rotation():
    rng = np.random.default_rng(1)
    for _ in range(20):
        axis = _unit(rng.normal(size=3))
        deg = float(rng.uniform(-180, 180))
        m = Transform3.from_axis_angle(axis, deg).matrix
        R = m[:3, :3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
        assert np.allclose(R @ axis, axis, atol=1e-14)  # the axis is fixed
        # a vector normal to the axis turns by the angle, counter-clockwise about the axis (right-handed)
        n = _unit(np.cross(axis, [1.0, 0.0, 0.0] if abs(axis[0]) < 0.9 else [0.0, 1.0, 0.0]))
        r = R @ n
        assert abs(math.degrees(math.atan2(np.dot(np.cross(n, r), axis), np.dot(n, r))) - deg) < 1e-9
    # 90 degrees about Y takes +Z to +X (glam's convention)
    assert np.allclose(Transform3.from_axis_angle(Y, 90.0).map_vector([0, 0, 1]), [1, 0, 0], atol=1e-15)


def test_then_and_correction():
    a, b = Transform3.from_translation((1, 2, 3)), Transform3.from_axis_angle(Y, 90.0)
    p = np.array([1.0, 0.0, 0.0])
    assert np.allclose(a.then(b).map_point(p), b.map_point(a.map_point(p)))
    c = np.array([0.381, 0.2975, 0.2655])
    t = b.corrected(c)
    assert np.allclose(t.map_point(c), c, atol=1e-15)  # the centre stays where it is
    assert np.allclose(t.map_vector([0, 0, 1]), [1, 0, 0], atol=1e-15)
    # the engine's corrected identity is exactly the identity, so such an object carries no transform
    m = Transform3.IDENTITY.corrected(c).matrix
    assert np.array_equal(m, np.eye(4))


# ------------------------------------------------------------------------------------------ closed forms


def _slab(o, d, lo, hi):
    """Entry distance and outward normal of a ray into an axis box (slab method)."""
    inv = 1.0 / d
    t1, t2 = (lo - o) * inv, (hi - o) * inv
    tn, tf = np.minimum(t1, t2), np.maximum(t1, t2)
    t0, t_far = tn.max(), tf.min()
    if t0 > t_far or t_far < 0:
        return None
    axis = int(np.argmax(tn))
    n = np.zeros(3)
    n[axis] = -np.sign(d[axis])
    return t0, n


def test_rotated_box_hit_is_the_closed_form():
    sc, cam, _ = presets.cornell()
    box = sc.objects[6]
    M = object_matrix(box)
    Minv = np.linalg.inv(M)
    lo, hi = np.minimum(box.mesh.a, box.mesh.b), np.maximum(box.mesh.a, box.mesh.b)
    centre_w = M[:3, :3] @ ((lo + hi) / 2) + M[:3, 3]
    lone = Scene([SimpleObject(box.mesh, box.material, box.transform)], SimpleSkybox)
    ts = oracle.TwinScene(lone)
    rng = np.random.default_rng(7)
    n_hit = 0
    for _ in range(200):
        o = np.asarray(cam.pos) + rng.normal(size=3) * 0.3
        d = _unit(centre_w + rng.normal(size=3) * 0.15 - o)
        ol, dl = Minv[:3, :3] @ o + Minv[:3, 3], Minv[:3, :3] @ d
        dl = dl / np.linalg.norm(dl)  # Ray::new normalises
        ref = _slab(ol, dl, lo, hi)
        h = ts.trace(o, d, precision=64)
        if ref is None:
            assert h is None
            continue
        n_hit += 1
        t_l, n_l = ref
        p_w = M[:3, :3] @ (ol + dl * t_l) + M[:3, 3]
        n_w = _unit(M[:3, :3] @ n_l)
        assert h is not None
        assert abs(h["t"] - np.linalg.norm(p_w - o)) < 1e-12
        assert np.allclose(h["pos"], p_w, atol=1e-12)
        assert np.allclose(h["normal"], n_w, atol=1e-12) and np.allclose(h["ray_normal"], n_w, atol=1e-12)
        assert h["front_face"]
    assert n_hit > 100


def test_scaled_sphere_normal_is_mapped_by_the_linear_part():
    """outgoing_intersection maps normals by the linear part of the transform, then normalises; under a
    non-uniform scaling that is not the ellipsoid's surface normal, and the engine's choice is kept."""
    S = Transform3.from_scale((2.0, 1.0, 0.5)).then(Transform3.from_axis_angle(_unit((1, 1, 0)), 30.0))
    obj = SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LambertianMaterial(), S)
    M = object_matrix(obj)
    A3, tt = M[:3, :3], M[:3, 3]
    Ainv = np.linalg.inv(A3)
    ts = oracle.TwinScene(Scene([obj], SimpleSkybox))
    rng = np.random.default_rng(3)
    for _ in range(100):
        o = rng.normal(size=3) * 0.2 + np.array([0.0, 0.0, -6.0])
        d = _unit(rng.normal(size=3) * 0.25 + np.array([0.0, 0.0, 1.0]))
        # the world ray against the ellipsoid |A^-1 (x - t)| = 1: a quadratic in the world distance
        q0, q1 = Ainv @ (o - tt), Ainv @ d
        a, b, c = q1 @ q1, 2 * q0 @ q1, q0 @ q0 - 1
        disc = b * b - 4 * a * c
        h = ts.trace(o, d, precision=64)
        if disc < 0:
            assert h is None
            continue
        t = (-b - math.sqrt(disc)) / (2 * a)
        p = o + d * t
        n_obj = Ainv @ (p - tt)  # the object-space sphere normal
        want = _unit(A3 @ n_obj)
        true_normal = _unit(Ainv.T @ n_obj)
        assert abs(h["t"] - t) < 1e-10 * t
        assert np.allclose(h["normal"], want, atol=1e-10)
        if abs(want @ true_normal) < 1 - 1e-6:  # distinguishable from the inverse-transpose normal
            assert not np.allclose(h["normal"], true_normal, atol=1e-6)


def test_rigidly_transformed_triangle_matches_the_moved_triangle():
    v = np.array([[-1.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.0, 1.2, 0.1]])
    n = np.array([_unit([0, 0, 1]), _unit([0.1, 0, 1]), _unit([0, 0.2, 1])])
    T = Transform3.from_axis_angle(_unit((0.3, 1.0, 0.2)), 40.0).then_translate((0.2, -0.1, 0.3))
    a = SimpleObject(TriangleMesh(v, n), LambertianMaterial(), T, corrected=False)
    M = T.matrix
    b = SimpleObject(TriangleMesh(v @ M[:3, :3].T + M[:3, 3], n @ M[:3, :3].T), LambertianMaterial())
    ta, tb = oracle.TwinScene(Scene([a], SimpleSkybox)), oracle.TwinScene(Scene([b], SimpleSkybox))
    rng = np.random.default_rng(5)
    centroid = (v.mean(axis=0)) @ M[:3, :3].T + M[:3, 3]
    hits = differ = 0
    for _ in range(100):
        o = np.array([0.0, 0.5, -4.0]) + rng.normal(size=3) * 0.1
        d = _unit(centroid + rng.normal(size=3) * 0.4 - o)
        ha, hb = ta.trace(o, d), tb.trace(o, d)
        if ha is None or hb is None:
            differ += (ha is None) != (hb is None)  # only a ray through an edge may tell them apart
            continue
        hits += 1
        assert abs(ha["t"] - hb["t"]) < 1e-12
        assert np.allclose(ha["normal"], hb["normal"], atol=1e-12)
        assert ha["front_face"] == hb["front_face"]
    assert hits > 50 and differ <= 1


def test_closest_hit_uses_world_distance_whatever_the_order():
    """A sphere shrunk four times (object distances four times its world distances) in front of a wall: the
    shrunk sphere is the closest hit whichever object the traversal tests first."""
    small = SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LambertianMaterial((0.9, 0.1, 0.1)),
                         Transform3.from_scale(0.25))
    wall = SimpleObject(ParallelogramMesh((-5.0, -5.0, 10.0), (10.0, 0.0, 0.0), (0.0, 10.0, 0.0)),
                        LambertianMaterial((0.1, 0.9, 0.1)))
    for objs in ([small, wall], [wall, small]):
        ts = oracle.TwinScene(Scene(objs, SimpleSkybox))
        for prec in (32, 64):
            for mode in (oracle.BOX_REFERENCE, oracle.BOX_TIGHT):
                h = ts.trace([0.0, 0.0, -10.0], [0.0, 0.0, 1.0], precision=prec, box_mode=mode)
                assert abs(h["t"] - 9.75) < 1e-5 and h["material"] == objs.index(small)


def test_transformed_volumes_hit_inside_the_mapped_boundary():
    """Media under a transform: the free flight is drawn in object space, and the hit is reported on the world
    ray, at the world distance, inside the mapped boundary.  A box boundary never yields a hit, transformed or not:
    its test only finds the faces facing the ray origin (axis_box.rs:64-145), so the exit search after the entry
    face finds nothing (object/volumetric.rs:109-117)."""
    rot = Transform3.from_axis_angle(_unit((0.2, 1.0, 0.4)), 33.0)
    vol = VolumetricObject(SphereMesh((0.0, 0.0, 0.0), 1.0), IsotropicMaterial((0.8, 0.8, 0.8), 0.0), 3.0,
                           Transform3.from_scale((1.5, 0.7, 1.0)).then(rot))
    box = VolumetricObject(AxisBoxMesh.new_centred((0.0, 0.0, 0.0), (2.0, 1.0, 1.5)),
                           IsotropicMaterial((0.8, 0.8, 0.8), 0.0), 3.0, rot)
    M = object_matrix(vol)
    Minv = np.linalg.inv(M)
    ts, tb = oracle.TwinScene(Scene([vol], SimpleSkybox)), oracle.TwinScene(Scene([box], SimpleSkybox))
    rng = np.random.default_rng(11)
    seen = 0
    for _ in range(100):
        o = np.array([0.0, 0.0, -5.0]) + rng.normal(size=3) * 0.05
        d = _unit(rng.normal(size=3) * 0.15 + np.array([0.0, 0.0, 1.0]))
        assert tb.trace(o, d) is None and tb.trace(o + d * 5.0, d) is None
        h = ts.trace(o, d)
        if h is None:
            continue
        seen += 1
        p = h["pos"]
        assert np.allclose(p, o + d * h["t"], atol=1e-9)  # on the world ray, at the world distance
        q = Minv[:3, :3] @ p + Minv[:3, 3]
        assert q @ q <= 1 + 1e-9  # inside the mapped sphere
        assert abs(np.linalg.norm(h["normal"]) - 1) < 1e-12
    assert seen > 50


# ------------------------------------------------------------------------------------------ packer


@pytest.mark.parametrize("make", [presets.cornell, presets.transformed_mix], ids=["cornell", "transformed_mix"])
def test_packer_builds_the_oracle_trees(make):
    sc, _, _ = make()
    h, obj, loc, need = prepare_scene(sc)
    ts = oracle.TwinScene(sc)
    assert h == ts.tree_hash()
    o2, l2 = ts.dfs_order()
    assert np.array_equal(obj, o2) and np.array_equal(loc, l2)


def test_identity_matrix_is_no_transform():
    sc, _, _ = presets.mixed_primitives()
    h0, obj0, _, _ = prepare_scene(sc)
    for o in sc.objects:
        if not isinstance(o.mesh, (TriangleListMesh, RaymarchedIsosurfaceMesh)):
            o.transform = Transform3.IDENTITY
    sc._cache = None
    assert sc.desc().n_transforms > 0
    h1, obj1, _, _ = prepare_scene(sc)
    assert h0 == h1 and np.array_equal(obj0, obj1)


def _one(obj):
    return Scene([obj, SimpleObject(SphereMesh((0, -100, 0), 99), LambertianMaterial())], SimpleSkybox)


def test_invalid_and_unsupported_transforms_are_refused():
    singular = Transform3.from_scale((1.0, 0.0, 1.0))
    with pytest.raises(RayError) as e:
        prepare_scene(_one(SimpleObject(SphereMesh((0, 0, 0), 1), LambertianMaterial(), singular)))
    assert e.value.status == A.RAYNA_INVALID_ARGUMENT
    rot = Transform3.from_axis_angle(Y, 10.0)
    with pytest.raises(RayError) as e:
        prepare_scene(_one(SimpleObject(RaymarchedIsosurfaceMesh(SdfSphere((0, 0, 0), 1.0)), LambertianMaterial(),
                                        rot)))
    assert e.value.status == A.RAYNA_UNSUPPORTED
    tv = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], float)
    tn = np.array([[[0, 0, 1]] * 3], float)
    prepare_scene(_one(SimpleObject(TriangleListMesh(tv, tn), LambertianMaterial(), rot)))  # an instance
    sc = _one(SimpleObject(SphereMesh((0, 0, 0), 1), LambertianMaterial()))
    d = sc.desc()
    d.objects[0].transform = 1  # no such matrix
    with pytest.raises(RayError) as e:
        prepare_scene(sc)
    assert e.value.status == A.RAYNA_INVALID_ARGUMENT
