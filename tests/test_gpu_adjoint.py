"""GPU path-replay adjoint (rayna_adjoint) through the C ABI: against the oracle's single-precision adjoint,
which retraces the same paths bit for bit; against central finite differences of the double-precision
oracle (h = 1e-3, frozen RNG); the reuse of a rendered frame's radiance; frames of several path chunks; and
the torch.autograd binding with a short inverse-rendering run."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, presets

from .test_adjoint_oracle import fd_subset

pytestmark = pytest.mark.gpu


def _setup(width, height, samples, seed=4):
    scene, cam, opts, hidden = presets.inverse_spheres(width=width, height=height, samples=samples)
    opts = replace(opts, seed=seed)
    target = np.random.default_rng(5).random((height, width, 3)).astype(np.float32) * 0.5
    return scene, cam, opts, hidden, target


def test_adjoint_matches_oracle_fp32(gpu):
    scene, cam, opts, hidden, target = _setup(64, 64, 8)
    print("case adjoint vs oracle fp32 64x64x8", flush=True)
    with Renderer(scene, cam, opts) as r:
        img = r.render(layout="rowmajor").img
        d = (2.0 * (img - target)).astype(np.float32)
        g_reuse, st = r.adjoint(d, frame=0, layout="rowmajor")          # primal = the rendered frame
        g_fresh, st2 = r.adjoint(d, frame=0, layout="rowmajor")         # primal recomputed
        assert st2.segments > st.segments
        # x-major input gives the same gradient
        g_x, _ = r.adjoint(np.ascontiguousarray(d.transpose(1, 0, 2)), frame=0, layout="xmajor")
    tw = oracle.TwinScene(scene)
    ref_img, _, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(ref_img, img)
    g32, _, _ = tw.adjoint(cam, opts, d, precision=32)
    for g in (g_reuse, g_fresh, g_x):
        rel = np.linalg.norm(g - g32) / np.linalg.norm(g32)
        assert rel <= 1e-9, rel
    assert np.array_equal(g32 != 0, g_reuse != 0)


def test_adjoint_matches_fp64_central_differences(gpu):
    """At the inverse-rendering configuration's size (256x256, 16 spp): the GPU computes in single precision,
    so a few of its paths part from the double-precision oracle's, which moves the gradient by about 1e-4
    (normwise) against central differences of the oracle (h = 1e-3, frozen RNG), inside the 1e-3 bound."""
    scene, cam, opts, hidden, target = _setup(256, 256, 16)
    print("case adjoint vs fp64 central differences 256x256x16", flush=True)
    with Renderer(scene, cam, opts) as r:
        img = r.render(layout="rowmajor").img
        d = (2.0 * (img.astype(np.float64) - target)).astype(np.float32)
        g, _ = r.adjoint(d, frame=0, layout="rowmajor")
    tw = oracle.TwinScene(scene)
    idx, fd = fd_subset(tw, cam, opts, scene.colours(), target.astype(np.float64), g, hidden)
    adj = np.array([g[hidden[i], c] for i, c in idx])
    rel = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    print("GPU adjoint vs fp64 central differences: normwise relative error", rel)
    assert rel <= 1e-3


def test_adjoint_frames_and_chunks(gpu):
    """A frame after the first (its own RNG key) and a frame of two path chunks (512x256 pixels x 72 spp
    with 2^23 paths per chunk) against the oracle."""
    scene, cam, opts, hidden, _ = _setup(512, 256, 72)
    print("case adjoint two chunks 512x256x72 frame 1", flush=True)
    d = np.random.default_rng(8).standard_normal((256, 512, 3)).astype(np.float32)
    with Renderer(scene, cam, opts) as r:
        r.set_tuning(chunk_slots=1 << 23)
        g, st = r.adjoint(d, frame=1, layout="rowmajor")
    tw = oracle.TwinScene(scene)
    g32, segs, _ = tw.adjoint(cam, opts, d, precision=32, frame=1, threads=16)
    assert st.segments == 2 * segs  # primal and replay
    rel = np.linalg.norm(g - g32) / np.linalg.norm(g32)
    assert rel <= 1e-9, rel


def test_set_material_colours_renders_like_a_rebuilt_scene(gpu):
    scene, cam, opts, hidden, _ = _setup(48, 32, 4)
    new = scene.colours() * np.float32(0.7)
    with Renderer(scene, cam, opts) as r:
        r.render()
        r.set_material_colours(new)
        img = r.render(layout="rowmajor").img
        assert r.render().stats.accum_frames == 2
    tw = oracle.TwinScene(scene)
    tw.set_colours(new)
    ref, _, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(ref, img)


def test_torch_function_and_short_inverse_run(gpu):
    import torch
    from rayna_amd.diff import ColourRenderer, recover_colours

    scene, cam, opts, hidden, target = _setup(64, 64, 8)
    f = ColourRenderer(scene, cam, opts)
    colours = torch.tensor(scene.colours(), requires_grad=True)
    img = f(colours)
    loss = ((img - torch.from_numpy(target)) ** 2).sum()
    loss.backward()
    with Renderer(scene, cam, opts) as r:
        ref = r.render(layout="rowmajor").img
        g, _ = r.adjoint(2.0 * (ref - target), layout="rowmajor")
    assert np.array_equal(img.detach().numpy(), ref)
    np.testing.assert_allclose(colours.grad.numpy(), g.astype(np.float32), rtol=1e-6, atol=1e-9)
    f.close()

    print("case inverse rendering 30 steps", flush=True)
    res = recover_colours(steps=30, target_samples=64)
    print("loss", res.losses[0], "->", res.losses[-1], "param rmse", res.param_rmse[0], "->", res.param_rmse[-1],
          "identifiable", res.identifiable, "adjoint/forward", res.adjoint_forward_ratio)
    assert res.losses[-1] < 0.7 * res.losses[0]
    assert res.param_rmse[-1] < res.param_rmse[0]
    assert res.identifiable > 300
