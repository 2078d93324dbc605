"""The oracle's path-replay adjoint (twin::Tracer::adjoint, twin::adjoint_frame), pinned by central finite
differences of the double-precision restatement with frozen RNG, and by closed forms: the radiance of a
convex lambertian object under WhiteSkybox is exactly its albedo (rayna_engine/tests/simple_renders.rs:86-141)
and a directly seen light is exactly its emission (light.rs:17-32)."""
import numpy as np
import pytest

import oracle
from rayna_amd import presets
from rayna_amd.scene import (Camera, LambertianMaterial, LightMaterial, NoSkybox, ParallelogramMesh, RenderOpts,
                             Scene, SimpleObject, SphereMesh, WhiteSkybox)

H = 1e-3  # central-difference step


def _loss_and_grad(tw, cam, opts, colours, target, precision=64):
    tw.set_colours(colours)
    img, _, _, _ = tw.render(cam, opts, precision=precision, box_mode=oracle.BOX_TIGHT)
    img = img.astype(np.float64)
    loss = float(((img - target) ** 2).sum())
    grad, _, _ = tw.adjoint(cam, opts, (2.0 * (img - target)).astype(np.float32), precision=precision)
    return loss, grad


def _fd(tw, cam, opts, colours, target, i, c):
    def loss(p):
        tw.set_colours(p)
        img, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT)
        return float(((img.astype(np.float64) - target) ** 2).sum())
    e = np.zeros_like(colours)
    e[i, c] = H
    out = (loss(colours + e) - loss(colours - e)) / (2 * H)
    tw.set_colours(colours)
    return out


def fd_subset(tw, cam, opts, colours, target, grad, hidden, n_top=12, n_rand=12, seed=0):
    """Central differences of the loss for the n_top hidden colours with the largest adjoint gradient and
    n_rand random identifiable ones: (indices, fd values)."""
    g = np.abs(grad[hidden])
    flat = np.argsort(-g.ravel())
    top = [divmod(int(k), 3) for k in flat[:n_top]]
    nz = [divmod(int(k), 3) for k in np.flatnonzero(g.ravel() > 0)]
    rng = np.random.default_rng(seed)
    pick = [nz[k] for k in rng.choice(len(nz), size=min(n_rand, len(nz)), replace=False)]
    idx = list(dict.fromkeys(top + pick))
    fd = np.array([_fd(tw, cam, opts, colours, target, int(hidden[i]), c) for i, c in idx])
    return idx, fd


@pytest.fixture(scope="module")
def inverse_small():
    scene, cam, opts, hidden = presets.inverse_spheres(width=40, height=40, samples=4)
    tw = oracle.TwinScene(scene)
    colours = scene.colours()
    target = np.random.default_rng(3).random((opts.height, opts.width, 3)) * 0.5
    return tw, cam, opts, hidden, colours, target


def test_adjoint_fp64_matches_central_differences(inverse_small):
    tw, cam, opts, hidden, colours, target = inverse_small
    _, grad = _loss_and_grad(tw, cam, opts, colours, target)
    idx, fd = fd_subset(tw, cam, opts, colours, target, grad, hidden)
    adj = np.array([grad[hidden[i], c] for i, c in idx])
    rel = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    print("fp64 adjoint vs central differences: normwise relative error", rel, "over", len(idx), "colours")
    assert rel <= 1e-4
    # the gradient of every colour that no path reaches is exactly zero; lights and spheres both appear
    assert (grad[hidden] != 0).any(axis=1).sum() > 20


def test_adjoint_fp32_close_to_fp64(inverse_small):
    tw, cam, opts, hidden, colours, target = inverse_small
    tw.set_colours(colours)
    img, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT)
    d = (2.0 * (img.astype(np.float64) - target)).astype(np.float32)
    g64, _, _ = tw.adjoint(cam, opts, d, precision=64)
    g32, _, _ = tw.adjoint(cam, opts, d, precision=32)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print("fp32 vs fp64 adjoint normwise relative difference", rel)
    assert rel <= 2e-3


def test_planted_sign_error_is_caught(inverse_small):
    """A plausible mistake, the light term with the wrong sign, fails the finite-difference check."""
    tw, cam, opts, hidden, colours, target = inverse_small
    _, grad = _loss_and_grad(tw, cam, opts, colours, target)
    idx, fd = fd_subset(tw, cam, opts, colours, target, grad, hidden)
    wrong = grad.copy()
    lights = [m for m, mat in enumerate(tw.scene.materials_list()) if isinstance(mat, LightMaterial)]
    wrong[lights] *= -1.0
    adj = np.array([wrong[hidden[i], c] for i, c in idx])
    assert np.linalg.norm(adj - fd) / np.linalg.norm(fd) > 1e-2


def test_albedo_gradient_closed_form():
    """Outside a convex lambertian sphere under WhiteSkybox every sample's radiance is the albedo (one bounce
    to the white sky) or 1 (sky), so with d_image = 1 the albedo gradient of channel c is the expected number
    of sphere samples per pixel, sum_px (1 - I_c) / (1 - albedo_c)."""
    albedo = (0.3, 0.55, 0.8)
    scene = Scene([SimpleObject(SphereMesh((0.0, 0.0, 0.0), 1.0), LambertianMaterial(albedo))], WhiteSkybox)
    cam = Camera(pos=(0.0, 0.0, 4.0), fwd=(0.0, 0.0, -1.0), v_fov=45.0, focus_dist=4.0)
    opts = RenderOpts(width=24, height=24, samples=5, ray_depth=4, seed=9)
    tw = oracle.TwinScene(scene)
    img, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT)
    grad, _, _ = tw.adjoint(cam, opts, np.ones((24, 24, 3), np.float32), precision=64)
    want = [(1.0 - img[..., c].astype(np.float64)).sum() / (1.0 - albedo[c]) for c in range(3)]
    np.testing.assert_allclose(grad[0], want, rtol=1e-5)
    assert want[0] > 24 * 24 * 0.2  # the sphere covers a good part of the frame


def test_emission_gradient_closed_form():
    """A light parallelogram seen directly against NoSkybox: each sample's radiance is the emission (T = 1)
    or 0, so with d_image = 1 the emission gradient is sum_px I_c / emission_c."""
    emission = (2.0, 3.0, 0.5)
    scene = Scene([SimpleObject(ParallelogramMesh((-1.0, -1.0, 0.0), (2.0, 0.0, 0.0), (0.0, 2.0, 0.0)),
                                LightMaterial(emission))], NoSkybox)
    cam = Camera(pos=(0.0, 0.0, 4.0), fwd=(0.0, 0.0, -1.0), v_fov=45.0, focus_dist=4.0)
    opts = RenderOpts(width=20, height=20, samples=4, ray_depth=3, seed=11)
    tw = oracle.TwinScene(scene)
    img, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT)
    grad, _, _ = tw.adjoint(cam, opts, np.ones((20, 20, 3), np.float32), precision=64)
    want = [img[..., c].astype(np.float64).sum() / emission[c] for c in range(3)]
    np.testing.assert_allclose(grad[0], want, rtol=1e-6)
    assert want[0] > 20
