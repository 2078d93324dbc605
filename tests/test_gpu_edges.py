"""Edge cases on the GPU, bit for bit against the oracle: an empty scene (sky only), a scene of unbounded SDFs
alone, a single triangle, an empty BVH mesh, a 1x1 image, ray depth 0, and all of them in one batch."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer
from rayna_amd.batch import BatchRenderer, Request
from rayna_amd.scene import (BvhMesh, Camera, LambertianMaterial, LightMaterial, MetalMaterial,
                             RaymarchedIsosurfaceMesh, RenderOpts, Scene, SdfSminSpheres, SimpleObject,
                             SimpleSkybox, SphereMesh, TriangleListMesh, WhiteSkybox)

pytestmark = pytest.mark.gpu

CAM = Camera(pos=(0.0, 1.0, 4.0), fwd=(0.0, -0.2, -1.0), v_fov=50.0, focus_dist=4.0)
OPTS = RenderOpts(width=24, height=16, samples=3, ray_depth=4, seed=7)
GROUND = SimpleObject(SphereMesh((0.0, -100.5, 0.0), 100.0), LambertianMaterial((0.5, 0.5, 0.5)))


def _tri():
    v = np.array([[[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.5, 0.0]]])
    n = np.tile(np.array([0.0, 0.0, 1.0]), (1, 3, 1))
    return TriangleListMesh(v, n)


CASES = {
    "empty_scene": lambda: Scene([], SimpleSkybox),
    "empty_scene_white": lambda: Scene([], WhiteSkybox),
    "sdf_only": lambda: Scene([SimpleObject(RaymarchedIsosurfaceMesh(
        SdfSminSpheres(np.array([[0.0, 0.3, 0.0], [0.6, 0.2, 0.0]]), np.array([0.5, 0.3]), 0.1), 128, 1e-4),
        MetalMaterial((0.8, 0.7, 0.6), 0.2))], SimpleSkybox),
    "single_triangle": lambda: Scene([SimpleObject(_tri(), LambertianMaterial((0.7, 0.3, 0.3))), GROUND],
                                     SimpleSkybox),
    "empty_bvh_mesh": lambda: Scene([SimpleObject(BvhMesh([]), LambertianMaterial()), GROUND,
                                     SimpleObject(SphereMesh((0.0, 0.2, 0.0), 0.4), LightMaterial((4.0, 3.0, 2.0)))],
                                    SimpleSkybox),
}


def _check(sc, opts):
    ref, segs, _, _ = oracle.TwinScene(sc).render(CAM, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    with Renderer(sc, CAM, opts) as r:
        out = r.render(layout="rowmajor")
    assert np.array_equal(out.img, ref)
    assert out.stats.segments == segs
    return ref


@pytest.mark.parametrize("name", sorted(CASES))
def test_edge_scene(gpu, name):
    sc = CASES[name]()
    print("case", name, flush=True)
    img = _check(sc, OPTS)
    if name.startswith("empty_scene"):
        assert np.ptp(img.reshape(-1, 3), axis=0).max() > 0 or name.endswith("white")  # the sky, nothing else
    _check(sc, replace(OPTS, width=1, height=1, samples=5))
    _check(sc, replace(OPTS, ray_depth=0))


def test_edge_scenes_in_one_batch(gpu):
    names = sorted(CASES)
    scenes = [CASES[n]() for n in names]
    reqs = [Request(i, CAM, replace(OPTS, seed=11 + i)) for i in range(len(names))]
    reqs.append(Request(0, CAM, replace(OPTS, width=1, height=1)))
    print("case edge batch", flush=True)
    with BatchRenderer(scenes) as b:
        b.set_tuning(chunk_slots=1 << 16, pipelines=2)
        imgs, _ = b.render(reqs)
    for q, img in zip(reqs, imgs):
        ref, _, _, _ = oracle.TwinScene(scenes[q.scene]).render(q.camera, q.opts, precision=32,
                                                               box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(img, ref), names[q.scene]
