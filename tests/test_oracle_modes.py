"""The oracle's non-PBR render modes (render_px_once, rayna_engine/src/render/renderer.rs:410-490) and the UI's
8-bit conversion (rayna_ui/src/ext/img_ext.rs:15-45), pinned by closed forms."""
import math

import numpy as np
import pytest

import oracle
from rayna_amd.scene import (AxisBoxMesh, Camera, LambertianMaterial, NoSkybox, ParallelogramMesh, RenderMode,
                             RenderOpts, Scene, SimpleObject, SphereMesh, WhiteSkybox)

PAL = np.array([[1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [1.0, 0.5, 0.0], [1.0, 1.0, 0.0], [0.5, 1.0, 0.0], [0.0, 1.0, 0.0],
                [0.0, 1.0, 0.5], [0.0, 1.0, 1.0], [0.0, 0.5, 1.0], [0.0, 0.0, 1.0], [0.5, 0.0, 1.0], [1.0, 0.0, 1.0],
                [0.0, 0.0, 0.0]], np.float32)


def _render(scene, cam, opts, precision=64):
    img, _, _, _ = oracle.TwinScene(scene).render(cam, opts, precision=precision, box_mode=oracle.BOX_TIGHT)
    return img


def _sphere_scene():
    return Scene([SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LambertianMaterial((0.3, 0.5, 0.7)))], WhiteSkybox)


def test_front_face_and_normals_of_a_sphere():
    cam = Camera(pos=(0.0, 0.0, -4.0), fwd=(0.0, 0.0, 1.0), v_fov=40.0, focus_dist=4.0)
    opts = RenderOpts(width=32, height=32, samples=1, ray_depth=3, mode=int(RenderMode.FrontFace))
    img = _render(_sphere_scene(), cam, opts)
    red = np.all(img == PAL[1], axis=-1)
    white = np.all(img == PAL[0], axis=-1)
    assert red.any() and white.any() and (red | white).all()  # front faces red, sky white
    # inside the sphere every face is a back face: the palette's white
    inside = _render(_sphere_scene(), Camera(pos=(0.0, 0.0, 0.0), fwd=(0.0, 0.0, 1.0), v_fov=40.0, focus_dist=1.0),
                     opts)
    assert np.all(inside == PAL[0])
    # outward normal at the centre pixel faces the camera: (0, 0, -1) -> (0.5, 0.5, 0)
    on = _render(_sphere_scene(), cam, RenderOpts(width=33, height=33, samples=1, ray_depth=3,
                                                  mode=int(RenderMode.OutwardNormal)))
    # (the sample region is offset by half a pixel, render_px_msaa, so the centre sample is slightly off axis)
    np.testing.assert_allclose(on[16, 16], [0.5, 0.5, 0.0], atol=0.03)
    rn = _render(_sphere_scene(), Camera(pos=(0.0, 0.0, 0.0), fwd=(0.0, 0.0, 1.0), v_fov=40.0, focus_dist=1.0),
                 RenderOpts(width=33, height=33, samples=1, ray_depth=3, mode=int(RenderMode.RayNormal)))
    np.testing.assert_allclose(rn[16, 16], [0.5, 0.5, 0.0], atol=0.03)  # from inside, the ray normal is -z


def test_distance_palette_ramp():
    """A wall at distance d fills the frame's centre: 2 cbrt(d) indexes the palette, lerping between entries."""
    wall = Scene([SimpleObject(ParallelogramMesh((-50.0, -50.0, 0.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0)),
                               LambertianMaterial((0.5, 0.5, 0.5)))], NoSkybox)
    for d, want in ((1.0, PAL[2]), (8.0, PAL[4]), (27.0, PAL[6]), (1.0 / 8.0, PAL[1]), (3.375, PAL[3])):
        cam = Camera(pos=(0.0, 0.0, -d), fwd=(0.0, 0.0, 1.0), v_fov=1.0, focus_dist=1.0)
        img = _render(wall, cam, RenderOpts(width=3, height=3, samples=1, ray_depth=1, mode=int(RenderMode.Distance)))
        np.testing.assert_allclose(img[1, 1], want, atol=1e-5)
    # halfway between palette entries 2 and 3 (d = 1.25^3)
    cam = Camera(pos=(0.0, 0.0, -1.953125), fwd=(0.0, 0.0, 1.0), v_fov=1.0, focus_dist=1.0)
    img = _render(wall, cam, RenderOpts(width=3, height=3, samples=1, ray_depth=1, mode=int(RenderMode.Distance)))
    np.testing.assert_allclose(img[1, 1], (PAL[2] + PAL[3]) / 2, atol=1e-5)


def test_uv_of_a_parallelogram_is_its_plane_coordinates():
    """The viewport plane (focus distance 1) is the quad's plane z = 0, so a sample at screen (cx, cy) hits
    world (-(cx - 4) / 4, -(cy - 4) / 4) and uv = (alpha, beta) = (0.5 - (cx - 4)/8, 0.5 - (cy - 4)/8); the mean
    over a pixel's 32 x 32 stratified samples is that at its sample region's centre x + 0.5 - 0.5/32 (the region
    is offset by half a stratum, renderer.rs:370-389)."""
    quad = Scene([SimpleObject(ParallelogramMesh((-1.0, -1.0, 0.0), (2.0, 0.0, 0.0), (0.0, 2.0, 0.0)),
                               LambertianMaterial((0.5, 0.5, 0.5)))], NoSkybox)
    cam = Camera(pos=(0.0, 0.0, -1.0), fwd=(0.0, 0.0, 1.0), v_fov=90.0, focus_dist=1.0)
    img = _render(quad, cam, RenderOpts(width=8, height=8, samples=1024, ray_depth=1, mode=int(RenderMode.Uv)))
    assert np.all(img[..., 2] == 0)
    c = np.arange(8) + 0.5 - 0.5 / 32
    want = 0.5 - (c - 4) / 8.0
    for y in range(1, 7):
        np.testing.assert_allclose(img[y, 1:7, 0], want[1:7], atol=2e-3)
        np.testing.assert_allclose(img[y, 1:7, 1], want[y], atol=2e-3)


def test_side_of_box_faces():
    """Box sides: 2 axis + (sgn > 0) with sgn = -sign(ray direction) (axis_box.rs:129), palette entry side % 12 + 1."""
    box = Scene([SimpleObject(AxisBoxMesh((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)), LambertianMaterial((0.5, 0.5, 0.5)))],
                NoSkybox)
    # seen from -z the ray runs along +z, so sgn.z = -1 and the side is 2 * 2 + 0 = 4
    views = {(0.0, 0.0, -5.0): 4, (0.0, 0.0, 5.0): 5, (-5.0, 0.0, 0.0): 0, (5.0, 0.0, 0.0): 1,
             (0.0, -5.0, 0.001): 2, (0.0, 5.0, 0.001): 3}
    for pos, side in views.items():
        fwd = tuple(-np.array(pos) / np.linalg.norm(pos))
        cam = Camera(pos=pos, fwd=fwd, v_fov=5.0, focus_dist=1.0)
        img = _render(box, cam, RenderOpts(width=3, height=3, samples=1, ray_depth=1, mode=int(RenderMode.Side)))
        assert np.array_equal(img[1, 1], PAL[side % 12 + 1]), (pos, img[1, 1])


def test_scatter_mode_is_the_scatter_direction():
    cam = Camera(pos=(0.0, 0.0, -4.0), fwd=(0.0, 0.0, 1.0), v_fov=40.0, focus_dist=4.0)
    img = _render(_sphere_scene(), cam, RenderOpts(width=24, height=24, samples=1, ray_depth=3,
                                                    mode=int(RenderMode.Scatter)))
    d = (img.astype(np.float64) - 0.5) * 2.0
    hit = ~np.all(img == 1.0, axis=-1)
    np.testing.assert_allclose(np.linalg.norm(d[hit], axis=-1), 1.0, atol=1e-5)  # unit scatter directions


def test_srgb8_matches_the_formula():
    rng = np.random.default_rng(0)
    img = np.concatenate([rng.random((64, 32, 3)), rng.random((64, 32, 3)) * 1.5,
                          np.zeros((2, 32, 3)), np.ones((2, 32, 3))]).astype(np.float32)
    img[0, 0] = [np.nan, -0.5, np.inf]
    out = oracle.to_srgb8(img)
    assert out.shape == (32, img.shape[0], 4) and np.all(out[..., 3] == 255)
    with np.errstate(invalid="ignore"):
        g = np.power(img.astype(np.float64), 1.0 / 2.2) * 255.0
    want = np.where(np.isnan(g) | (g <= 0), 0, np.where(g >= 255, 255, np.floor(np.nan_to_num(g)))).astype(np.int64)
    got = out[..., :3].transpose(1, 0, 2).astype(np.int64)
    diff = np.abs(got - want)
    assert diff.max() <= 1 and (diff > 0).mean() < 2e-3  # f32 pow vs double: truncation boundaries only
    assert list(out[0, 0, :3]) == [0, 0, 255]
    assert np.all(out[:, -1, :3] == 255) and np.all(out[:, -3, :3] == 0)
