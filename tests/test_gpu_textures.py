"""Textures and the image sky on the GPU, bit for bit against the oracle's fp32 restatement: a scene with every
texture kind (world and uv checkers, nested checkers, an image, uv / world / local noise, an image-textured light)
under an image sky -- images, a non-PBR mode, the adjoint (textured colours receive no gradient) and a batch
mixing textured and plain scenes."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, RenderMode, presets
from rayna_amd.batch import BatchRenderer, Request
from rayna_amd.scene import LambertianMaterial, MetalMaterial, SimpleObject, SphereMesh

pytestmark = pytest.mark.gpu


def test_textured_scene_bit_exact(gpu):
    sc, cam, opts = presets.textured(width=96, height=64, samples=4)
    print("case textured", opts, flush=True)
    tw = oracle.TwinScene(sc)
    with Renderer(sc, cam, opts) as r:
        assert r.tree_hash() == tw.tree_hash()
        out = r.render(layout="rowmajor")
        r.clear_accumulation()
        x = r.render(layout="xmajor").img
    ref, segs, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(out.img, ref)
    assert np.array_equal(x, ref.transpose(1, 0, 2))
    assert out.stats.segments == segs


def test_textured_mode_sees_the_image_sky(gpu):
    sc, cam, opts = presets.textured(width=64, height=48, samples=2)
    for mode in (RenderMode.Uv, RenderMode.OutwardNormal):
        o = replace(opts, mode=int(mode))
        print("case textured mode", mode.name, flush=True)
        with Renderer(sc, cam, o) as r:
            img = r.render(layout="rowmajor").img
        ref, _, _, _ = oracle.TwinScene(sc).render(cam, o, precision=32, box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(img, ref)


def test_textured_adjoint(gpu):
    sc, cam, opts = presets.textured(width=64, height=48, samples=4)
    sc.objects.append(SimpleObject(SphereMesh((0.7, 0.35, 2.4), 0.35), LambertianMaterial((0.7, 0.6, 0.2))))
    sc.objects.append(SimpleObject(SphereMesh((-0.8, 0.3, 2.2), 0.3), MetalMaterial((0.6, 0.7, 0.8), 0.1)))
    sc._cache = None
    print("case textured adjoint", flush=True)
    target = np.random.default_rng(11).random((opts.height, opts.width, 3)).astype(np.float32) * 0.5
    with Renderer(sc, cam, opts) as r:
        img = r.render(layout="rowmajor").img
        d = (2.0 * (img - target)).astype(np.float32)
        g, _ = r.adjoint(d, frame=0, layout="rowmajor")
    g32, _, _ = oracle.TwinScene(sc).adjoint(cam, opts, d, precision=32)
    textured = np.array([sc.desc().materials[i].texture > 0 for i in range(sc.desc().n_materials)])
    assert np.all(g[textured] == 0) and np.all(g32[textured] == 0)
    assert np.linalg.norm(g32) > 0
    assert np.linalg.norm(g - g32) / np.linalg.norm(g32) <= 1e-9


def test_textured_scenes_in_a_batch(gpu):
    scenes = [presets.textured()[0], presets.cornell()[0], presets.mixed_primitives()[0]]
    t = presets.textured(width=48, height=32, samples=3)
    c = presets.cornell(width=32, height=32, samples=4)
    m = presets.mixed_primitives(width=40, height=24, samples=2)
    reqs = [Request(0, t[1], t[2]), Request(1, c[1], c[2]), Request(2, m[1], m[2]),
            Request(0, t[1], replace(t[2], seed=9), layout="xmajor")]
    print("case textured batch", flush=True)
    with BatchRenderer(scenes) as b:
        imgs, _ = b.render(reqs)
    for q, img in zip(reqs, imgs):
        ref, _, _, _ = oracle.TwinScene(scenes[q.scene]).render(q.camera, q.opts, precision=32,
                                                               box_mode=oracle.BOX_TIGHT)
        if q.layout == "xmajor":
            ref = ref.transpose(1, 0, 2)
        assert np.array_equal(img, ref)
