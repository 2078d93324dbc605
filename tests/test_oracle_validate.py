"""The oracle's restatement of the engine's debug invariants (rayna_engine/src/shared/validate.rs:13-198, asserted
along every path at render/renderer.rs:395-561), counted instead of asserted.

Valid scenes break none of them in either precision; scenes that break one on purpose give the counts their
closed forms fix (every sample's path is known: a light filling the view, or a negative albedo under a white
sky).  The GPU's validate mode is checked against these counts in tests/test_gpu_validate.py.
"""
import numpy as np
import pytest

import oracle
from rayna_amd import (Camera, LambertianMaterial, LightMaterial, ParallelogramMesh, RenderOpts, Scene,
                       SimpleObject, SphereMesh, WhiteSkybox, NoSkybox, presets)

ZERO = dict.fromkeys(oracle.CHECKS, 0)


def light_wall(emission, w=16, h=12, spp=3):
    """A light quad filling the whole view: every camera ray hits it at depth 0 and the path ends there."""
    scene = Scene([SimpleObject(ParallelogramMesh((-50.0, -50.0, 5.0), (100.0, 0.0, 0.0), (0.0, 100.0, 0.0)),
                                LightMaterial(tuple(emission)))], NoSkybox)
    cam = Camera(pos=(0.0, 0.0, 0.0), fwd=(0.0, 0.0, 1.0), v_fov=40.0, focus_dist=1.0)
    return scene, cam, RenderOpts(width=w, height=h, samples=spp, ray_depth=4, seed=3)


def negative_ball(w=16, h=16, spp=2):
    """A Lambertian sphere of albedo (-0.5, 0.5, 0.5) filling the view under a white sky (convex: every
    scattered ray escapes to the sky after one bounce)."""
    return presets.sphere_colours((-0.5, 0.5, 0.5), -1.01, w, h, spp)


CASES = {
    "random_spheres": lambda: presets.random_spheres(width=32, height=18, samples=4),
    "mixed_primitives": lambda: presets.mixed_primitives(width=32, height=20, samples=4),
    "icosphere_mesh": lambda: presets.triangle_mesh(subdiv=3, width=32, height=18, samples=4),
    "sdf_volume": lambda: presets.sdf_volume(width=24, height=14, samples=2),
    "cornell": lambda: presets.cornell(width=24, height=24, samples=4),
    "rttnw": lambda: presets.rttnw(width=24, height=24, samples=4),
    "textured": lambda: presets.textured(width=24, height=16, samples=2),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("precision", [32, 64])
def test_valid_scenes_break_no_invariant(name, precision):
    sc, cam, opts = CASES[name]()
    _, counts = oracle.TwinScene(sc).render_checked(cam, opts, precision)
    assert counts == ZERO, counts


@pytest.mark.parametrize("precision", [32, 64])
def test_negative_emission_closed_form(precision):
    """Per sample: the light's colour and its emitted radiance T * Le break validate::colour (renderer.rs:530),
    the path ends at the light with that radiance (renderer.rs:395), and every pixel's mean is negative
    (renderer.rs:404)."""
    sc, cam, opts = light_wall((-1.0, 0.5, 0.5))
    img, counts = oracle.TwinScene(sc).render_checked(cam, opts, precision)
    n_px = opts.width * opts.height
    n = n_px * opts.samples
    assert counts == dict(ZERO, colour=2 * n, sample=n, pixel=n_px), counts
    assert np.all(img[..., 0] == -1.0)
    _, ok = oracle.TwinScene(light_wall((1.0, 0.5, 0.5))[0]).render_checked(cam, opts, precision)
    assert ok == ZERO


@pytest.mark.parametrize("precision", [32, 64])
def test_negative_albedo_closed_form(precision):
    """Per sample: the material colour, the throughput after the Lambertian bounce, the sky's radiance T * 1
    and the sample's radiance are negative in red; the pixel mean too."""
    sc, cam, opts = negative_ball()
    img, counts = oracle.TwinScene(sc).render_checked(cam, opts, precision)
    n_px = opts.width * opts.height
    n = n_px * opts.samples
    assert counts == dict(ZERO, colour=3 * n, sample=n, pixel=n_px), counts
    assert np.all(img[..., 0] == -0.5) and np.all(img[..., 1] == 0.5)
