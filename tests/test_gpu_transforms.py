"""Objects with a non-identity ObjectTransform (rayna_engine/src/object/transform.rs) and BVH meshes
(mesh/advanced/bvh.rs) on the GPU, bit for bit against the oracle's fp32 restatement: the engine's CORNELL
preset (two boxes rotated about their centres), a scene with every transformable kind under rotations, scalings
and shears, the engine's RTTNW preset (a BvhMesh floor and an instanced cube of balls) and an instanced triangle
list -- primary hits, images, the non-PBR modes, the adjoint and batched rendering."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, RenderMode, presets
from rayna_amd.batch import BatchRenderer, Request
from rayna_amd.scene import (Camera, LambertianMaterial, MetalMaterial, RenderOpts, Scene, SimpleObject, SimpleSkybox,
                             SphereMesh, Transform3, TriangleListMesh)

pytestmark = pytest.mark.gpu


def _instanced_mesh(width=64, height=48, samples=4):
    tv, tn, _ = presets.icosphere(3, seed=1)
    rot = Transform3.from_axis_angle(np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5]), 70.0)
    objs = [SimpleObject(TriangleListMesh(tv, tn), MetalMaterial((0.8, 0.7, 0.6), 0.2),
                         Transform3.from_scale((1.0, 0.6, 1.4)).then(rot).then_translate((-1.2, 1.0, 0.0))),
            SimpleObject(TriangleListMesh(tv, tn), LambertianMaterial((0.3, 0.6, 0.4)),
                         rot.then_translate((1.2, 1.0, 0.3))),
            SimpleObject(SphereMesh((0.0, -1000.0, 0.0), 1000.0), LambertianMaterial((0.5, 0.5, 0.5)))]
    cam = Camera(pos=(0.0, 1.5, 5.0), fwd=(0.0, -0.1, -1.0), v_fov=45.0, focus_dist=5.0)
    return Scene(objs, SimpleSkybox), cam, RenderOpts(width=width, height=height, samples=samples, ray_depth=6,
                                                      seed=3)


SCENES = {
    "cornell": lambda: presets.cornell(width=64, height=64, samples=8),
    "transformed_mix": lambda: presets.transformed_mix(width=96, height=64, samples=4),
    "rttnw": lambda: presets.rttnw(width=64, height=64, samples=4),
    "instanced_mesh": _instanced_mesh,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_transformed_scene_bit_exact(gpu, name):
    sc, cam, opts = SCENES[name]()
    print("case", name, opts, flush=True)
    tw = oracle.TwinScene(sc)
    with Renderer(sc, cam, opts) as r:
        assert r.tree_hash() == tw.tree_hash()
        rank, t = r.primary_hits(0)
        out = r.render(layout="rowmajor")
    rank32, t32 = tw.primary_hits(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, sample=0)
    assert np.array_equal(rank, rank32)
    hit = rank != -1
    assert hit.mean() > 0.3
    assert np.array_equal(t[hit], t32[hit].astype(np.float32))
    ref, segs, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(out.img, ref)
    assert out.stats.segments == segs


MODES = [RenderMode.RayNormal, RenderMode.OutwardNormal, RenderMode.Scatter, RenderMode.FrontFace,
         RenderMode.Distance, RenderMode.Uv, RenderMode.Side]


@pytest.mark.parametrize("mode", MODES, ids=[m.name for m in MODES])
def test_transformed_modes(gpu, mode):
    for name, (sc, cam, opts) in (("transformed_mix", presets.transformed_mix(width=96, height=64, samples=2)),
                                  ("rttnw", presets.rttnw(width=48, height=48, samples=2))):
        opts = replace(opts, mode=int(mode))
        print("case transformed mode", name, mode.name, flush=True)
        with Renderer(sc, cam, opts) as r:
            img = r.render(layout="rowmajor").img
        ref, _, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(img, ref)


def test_transformed_adjoint(gpu):
    sc, cam, opts = presets.transformed_mix(width=64, height=48, samples=4)
    print("case transformed adjoint", flush=True)
    target = np.random.default_rng(9).random((opts.height, opts.width, 3)).astype(np.float32) * 0.5
    with Renderer(sc, cam, opts) as r:
        img = r.render(layout="rowmajor").img
        d = (2.0 * (img - target)).astype(np.float32)
        g, _ = r.adjoint(d, frame=0, layout="rowmajor")
    g32, _, _ = oracle.TwinScene(sc).adjoint(cam, opts, d, precision=32)
    assert np.linalg.norm(g32) > 0
    assert np.linalg.norm(g - g32) / np.linalg.norm(g32) <= 1e-9


def test_transformed_scenes_in_a_batch(gpu):
    scenes = [presets.cornell()[0], presets.transformed_mix()[0], presets.rttnw()[0]]
    c0 = presets.cornell(width=48, height=40, samples=4)
    c1 = presets.transformed_mix(width=40, height=32, samples=3)
    c2 = presets.rttnw(width=40, height=40, samples=3)
    reqs = [Request(0, c0[1], c0[2]), Request(1, c1[1], c1[2]), Request(2, c2[1], c2[2]),
            Request(0, c0[1], replace(c0[2], seed=7))]
    print("case transformed batch", flush=True)
    with BatchRenderer(scenes) as b:
        imgs, _ = b.render(reqs)
    for q, img in zip(reqs, imgs):
        ref, _, _, _ = oracle.TwinScene(scenes[q.scene]).render(q.camera, q.opts, precision=32,
                                                               box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(img, ref)
