"""The non-PBR render modes and the UI's 8-bit output on the GPU, bit for bit against the oracle's fp32
restatement, on a scene with every primitive kind (spheres, triangles, a triangle list, a parallelogram light,
boxes, sphere and box media, a smooth-min SDF)."""
import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, RenderMode, presets
from dataclasses import replace

pytestmark = pytest.mark.gpu

MODES = [RenderMode.RayNormal, RenderMode.OutwardNormal, RenderMode.Scatter, RenderMode.FrontFace,
         RenderMode.Distance, RenderMode.Uv, RenderMode.Side]


@pytest.mark.parametrize("mode", MODES, ids=[m.name for m in MODES])
def test_mode_matches_oracle(gpu, mode):
    sc, cam, opts = presets.mixed_primitives(width=96, height=64, samples=4)
    opts = replace(opts, mode=int(mode))
    print("case mode", mode.name, flush=True)
    with Renderer(sc, cam, opts) as r:
        out = r.render(layout="rowmajor")
    ref, segs, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(out.img, ref)
    assert out.stats.segments == segs == opts.width * opts.height * opts.samples


def test_sdf_side_is_the_march_iteration(gpu):
    sc, cam, opts = presets.sdf_volume(width=64, height=36, samples=2)
    opts = replace(opts, mode=int(RenderMode.Side))
    with Renderer(sc, cam, opts) as r:
        img = r.render(layout="rowmajor").img
    ref, _, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    assert np.array_equal(img, ref)


def test_srgb8_output_matches_oracle(gpu):
    sc, cam, opts = presets.mixed_primitives(width=96, height=64, samples=4)
    with Renderer(sc, cam, opts) as r:
        rgba = r.render_srgb8().img
        r.clear_accumulation()
        img = r.render(layout="xmajor").img
    assert rgba.shape == (64, 96, 4)
    assert np.array_equal(rgba, oracle.to_srgb8(img))
    bad = replace(cam, v_fov=0.0)
    with Renderer(sc, bad, opts) as r:
        chk = r.render_srgb8().img
    assert chk.shape == (6, 6, 4)  # ilog2(64) x ilog2(96)
    assert set(np.unique(chk[..., :3])) <= {0, 255}
