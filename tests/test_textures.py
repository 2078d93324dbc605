"""Textures (rayna_engine/src/texture/*.rs) and the image skybox (skybox/hdri.rs) in the oracle.

Pinned by closed forms computed here: checker parity from the floors of the scaled, offset coordinates; bilinear
image lookups against a direct numpy evaluation of the clamped-corner rule, and exact texels at integer
coordinates; the improved Perlin noise through properties that a wrong fade, gradient table or permutation would
break (zero on the lattice, lattice gradients among the twelve cube-edge directions, period 256, the permutation a
permutation); the image sky against an emissive image-textured unit sphere seen from inside (both look up the
same sphere uv), the sphere uv itself being pinned by the render-mode tests.  Perlin parity with the engine's
`noise` crate is unpinned: that crate is not vendored and draws its own permutation.
"""
import numpy as np
import pytest

import oracle
from rayna_amd import presets
from rayna_amd.scene import (Camera, ColourSource, HdrImageSkybox, ImageTexture, LambertianMaterial, LightMaterial,
                             NoSkybox, Perlin, RenderOpts, Scene, ScalePoint, SimpleObject, SphereMesh,
                             UvCheckerTexture, UvNoiseTexture, WorldCheckerTexture, WorldNoiseTexture)


def _scene_with(*textures, sky=NoSkybox):
    objs = [SimpleObject(SphereMesh((0.0, 0.0, 5.0 * i), 1.0), LambertianMaterial(t)) for i, t in enumerate(textures)]
    return Scene(objs, sky)


def _tex_index(scene, i):
    """0-based texture index of object i's albedo (textures are listed before the material that uses them)."""
    return scene.desc().materials[i].texture - 1


@pytest.mark.parametrize("prec", [32, 64])
def test_world_and_uv_checkers(prec):
    even, odd = (0.9, 0.8, 0.7), (0.1, 0.2, 0.3)
    w = WorldCheckerTexture(even=even, odd=odd, scale=0.5, offset=(0.25, 0.0, 0.1))
    u = UvCheckerTexture(even=even, odd=odd, scale=0.25, offset=(0.1, 0.0))
    sc = _scene_with(w, u)
    ts = oracle.TwinScene(sc)
    rng = np.random.default_rng(1)
    for _ in range(300):
        p = rng.uniform(-4, 4, size=3)
        cells = np.floor(p / 0.5 + np.array([0.25, 0.0, 0.1]))
        want = even if int(cells.sum()) % 2 == 0 else odd
        assert np.array_equal(ts.texture_value(_tex_index(sc, 0), p, precision=prec), np.float32(want))
        uv = rng.uniform(0, 1, size=2)
        cells = np.floor(uv / 0.25 + np.array([0.1, 0.0]))
        want = even if int(cells.sum()) % 2 == 0 else odd
        assert np.array_equal(ts.texture_value(_tex_index(sc, 1), p, uv=uv, precision=prec), np.float32(want))


def test_nested_checker_walks_to_a_leaf():
    inner = WorldCheckerTexture(even=(1.0, 0.0, 0.0), odd=(0.0, 1.0, 0.0), scale=1.0)
    outer = UvCheckerTexture(even=inner, odd=(0.0, 0.0, 1.0), scale=0.5)
    sc = _scene_with(outer)
    ts = oracle.TwinScene(sc)
    t = _tex_index(sc, 0)
    assert np.array_equal(ts.texture_value(t, (0.5, 0.5, 0.5), uv=(0.75, 0.1)), np.float32([0, 0, 1]))  # odd cell
    assert np.array_equal(ts.texture_value(t, (0.5, 0.5, 0.5), uv=(0.25, 0.1)), np.float32([1, 0, 0]))  # inner even
    assert np.array_equal(ts.texture_value(t, (1.5, 0.5, 0.5), uv=(0.25, 0.1)), np.float32([0, 1, 0]))  # inner odd


def _bilinear(img, px, py):
    """Image::get_bilinear restated from its documentation: corners floor / ceil clamped to the image, the
    fraction measured from the clamped floor."""
    h, w, _ = img.shape
    fx, fy = np.floor(px), np.floor(py)
    x1, x2 = int(np.clip(fx, 0, w - 1)), int(np.clip(np.ceil(px), 0, w - 1))
    y1, y2 = int(np.clip(fy, 0, h - 1)), int(np.clip(np.ceil(py), 0, h - 1))
    xl, yl = px - np.clip(fx, 0, w - 1), py - np.clip(fy, 0, h - 1)
    c = img.astype(np.float64)
    top = c[y1, x1] + (c[y1, x2] - c[y1, x1]) * xl
    bottom = c[y2, x1] + (c[y2, x2] - c[y2, x1]) * xl
    return top + (bottom - top) * yl


def test_image_texture_is_bilinear_with_clamped_corners():
    img = presets.test_image(13, 9, 4)
    tex = ImageTexture(img, scale=(1.5, 0.75), offset=(-0.1, 0.2))
    sc = _scene_with(tex)
    ts = oracle.TwinScene(sc)
    t = _tex_index(sc, 0)
    rng = np.random.default_rng(2)
    for _ in range(300):
        uv = rng.uniform(-0.3, 1.3, size=2)
        u = -0.1 + uv[0] * 1.5
        v = 1.0 - (0.2 + uv[1] * 0.75)
        want = _bilinear(img, u * 13, v * 9)
        got = ts.texture_value(t, (0, 0, 0), uv=uv, precision=64)
        assert np.allclose(got, want, atol=1e-6)
    # integer coordinates read one texel exactly, in both precisions
    img = presets.test_image(16, 8, 5)  # power-of-two sides: u * width is exact in both precisions
    sc = _scene_with(ImageTexture(img))
    ts = oracle.TwinScene(sc)
    for x, y in ((0, 0), (15, 7), (5, 3), (7, 0)):
        uv = (x / 16.0, 1.0 - y / 8.0)
        for prec in (32, 64):
            assert np.array_equal(ts.texture_value(0, (0, 0, 0), uv=uv, precision=prec), img[y, x])


def test_noise_permutation_is_a_permutation():
    for seed in (0, 1, 69, 2**32 - 1):
        p = oracle.noise_permutation(seed)
        assert sorted(p[:256]) == list(range(256)) and np.array_equal(p[256:], p[:256])
    assert not np.array_equal(oracle.noise_permutation(1), oracle.noise_permutation(2))


EDGES = {(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if abs(a) + abs(b) + abs(c) == 2}


@pytest.mark.parametrize("prec", [32, 64])
def test_perlin_noise_properties(prec):
    sc = _scene_with(WorldNoiseTexture(ColourSource.greyscale(Perlin(69))))
    ts = oracle.TwinScene(sc)
    noise = lambda p: 2.0 * (float(ts.texture_value(0, p, precision=prec)[0]) - 0.5)
    rng = np.random.default_rng(3)
    for _ in range(50):
        c = rng.integers(-300, 300, size=3).astype(float)
        assert noise(c) == 0.0  # zero on the lattice
        if prec == 64:  # the gradient at a lattice point is one of the twelve cube-edge vectors
            e = 1e-3  # near a corner the noise is g . d up to O(d^3); its value passes through a float channel
            g = [(noise(c + e * np.eye(3)[k]) - noise(c - e * np.eye(3)[k])) / (2 * e) for k in range(3)]
            assert tuple(int(round(x)) for x in g) in EDGES and np.allclose(g, np.round(g), atol=1e-3)
        p = rng.uniform(-50, 50, size=3)
        assert abs(noise(p)) <= 1.1
        assert noise(p) == pytest.approx(noise(p + np.array([256.0, 0.0, 0.0])), abs=1e-4 if prec == 32 else 1e-9)
    # scaled and rgb sources: ScalePoint multiplies the point; three equal seeds give three equal channels
    scaled = _scene_with(WorldNoiseTexture(ColourSource.greyscale(ScalePoint(Perlin(69), 4.0))),
                         WorldNoiseTexture(ColourSource.rgb(Perlin(5), Perlin(5), Perlin(5))),
                         UvNoiseTexture(ColourSource.greyscale(Perlin(69))))
    t2 = oracle.TwinScene(scaled)
    for _ in range(20):
        p = rng.uniform(-3, 3, size=3)
        assert np.array_equal(t2.texture_value(0, p, precision=prec), ts.texture_value(0, p * 4.0, precision=prec))
        c = t2.texture_value(1, p, precision=prec)
        assert c[0] == c[1] == c[2]
        uv = rng.uniform(0, 1, size=2)
        assert np.array_equal(t2.texture_value(2, (9.0, 9.0, 9.0), uv=uv, precision=prec),
                              ts.texture_value(0, (uv[0], uv[1], 0.0), precision=prec))


def test_image_sky_matches_an_emissive_image_sphere():
    img = presets.test_image(40, 20, 9)
    tex = ImageTexture(img)
    inside = Scene([SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LightMaterial(tex))], NoSkybox)
    sky = Scene([SimpleObject(SphereMesh((0.0, 0.0, 50.0), 0.1), LambertianMaterial())], HdrImageSkybox(img))
    ti, ts = oracle.TwinScene(inside), oracle.TwinScene(sky)
    rng = np.random.default_rng(4)
    for prec in (32, 64):
        cam = Camera(pos=(0.0, 0.0, 0.0), fwd=(0.3, 0.2, -1.0), v_fov=120.0, focus_dist=1.0)
        opts = RenderOpts(width=24, height=16, samples=1, ray_depth=0, seed=3)
        a, _, _, _ = ti.render(cam, opts, precision=prec, box_mode=oracle.BOX_TIGHT)
        b, _, _, _ = ts.render(cam, opts, precision=prec, box_mode=oracle.BOX_TIGHT)
        assert np.allclose(a, b, atol=1e-5 if prec == 32 else 1e-12)
    for _ in range(50):  # a ray that misses everything sees the image at its direction's sphere uv
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        th, ph = np.arccos(-d[1]), np.arctan2(-d[2], d[0]) + np.pi
        want = _bilinear(img, ph / (2 * np.pi) * 40, (1 - th / np.pi) * 20)
        assert np.allclose(ts.sky(d), want, atol=1e-5)


def test_textured_scene_renders_in_both_precisions():
    sc, cam, opts = presets.textured(width=48, height=32, samples=2)
    ts = oracle.TwinScene(sc)
    a, _, _, _ = ts.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE)
    b, _, _, _ = ts.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    close = np.all(np.abs(a - b) <= 1e-3 * np.maximum(1.0, np.abs(a)), axis=-1)
    assert close.mean() > 0.9  # the rest: fp32 paths that part from fp64 at a branch
    assert np.isfinite(a).all() and np.isfinite(b).all()
