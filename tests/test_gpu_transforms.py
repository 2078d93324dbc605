"""Objects with a non-identity ObjectTransform (rayna_engine/src/object/transform.rs) on the GPU, bit for bit
against the oracle's fp32 restatement: the engine's CORNELL preset (two boxes rotated about their centres) and a
scene with every transformable kind under rotations, scalings and shears -- primary hits, images, the non-PBR
modes, the adjoint and batched rendering."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, RenderMode, presets
from rayna_amd.batch import BatchRenderer, Request

pytestmark = pytest.mark.gpu

SCENES = {
    "cornell": lambda: presets.cornell(width=64, height=64, samples=8),
    "transformed_mix": lambda: presets.transformed_mix(width=96, height=64, samples=4),
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
    assert hit.mean() > 0.5
    assert np.array_equal(t[hit], t32[hit].astype(np.float32))
    ref, segs, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(out.img, ref)
    assert out.stats.segments == segs


MODES = [RenderMode.RayNormal, RenderMode.OutwardNormal, RenderMode.Scatter, RenderMode.FrontFace,
         RenderMode.Distance, RenderMode.Uv, RenderMode.Side]


@pytest.mark.parametrize("mode", MODES, ids=[m.name for m in MODES])
def test_transformed_modes(gpu, mode):
    sc, cam, opts = presets.transformed_mix(width=96, height=64, samples=2)
    opts = replace(opts, mode=int(mode))
    print("case transformed mode", mode.name, flush=True)
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
    scenes = [presets.cornell()[0], presets.transformed_mix()[0]]
    c0 = presets.cornell(width=48, height=40, samples=4)
    c1 = presets.transformed_mix(width=40, height=32, samples=3)
    reqs = [Request(0, c0[1], c0[2]), Request(1, c1[1], c1[2]), Request(0, c0[1], replace(c0[2], seed=7))]
    print("case transformed batch", flush=True)
    with BatchRenderer(scenes) as b:
        imgs, _ = b.render(reqs)
    for q, img in zip(reqs, imgs):
        ref, _, _, _ = oracle.TwinScene(scenes[q.scene]).render(q.camera, q.opts, precision=32,
                                                               box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(img, ref)
