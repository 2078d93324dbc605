"""Camera rays traced as wave packets, on the GPU.

At depth 0 of a single-request frame whose chunk holds at least 16 samples per pixel, in a scene without
instances, the closest hits come from the packet traversal (k_extend_packet); chunks with fewer samples per pixel
take the per-lane traversal (k_extend).  Both must give the fp32 twin's image bit for bit, with the same number
of closest-hit queries, and the same image as each other at the benchmark's full size.
"""
import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, presets

pytestmark = pytest.mark.gpu

SCENES = {
    "random_spheres": lambda: presets.random_spheres(width=24, height=16, samples=16),
    "mixed_primitives": lambda: presets.mixed_primitives(width=24, height=16, samples=32),
    "triangle_mesh_s4": lambda: presets.triangle_mesh(subdiv=4, width=32, height=18, samples=64),
    "sdf_volume": lambda: presets.sdf_volume(width=16, height=9, samples=16),
    "cornell": lambda: presets.cornell(width=20, height=20, samples=16),
}


def _render_both(sc, cam, opts):
    """The frame traced in one chunk (packets at depth 0), then in chunks of 8 samples per pixel (per lane)."""
    with Renderer(sc, cam, opts) as r:
        r.clear_accumulation()
        a = r.render(layout="rowmajor")
        r.set_tuning(chunk_slots=opts.width * opts.height * 8)
        r.clear_accumulation()
        b = r.render(layout="rowmajor")
    assert b.stats.extend_launches > a.stats.extend_launches
    return a, b


@pytest.mark.parametrize("name", sorted(SCENES))
def test_packets_match_the_oracle_and_the_lane_traversal(gpu, name):
    sc, cam, opts = SCENES[name]()
    print("case packets", name, opts.width, opts.height, opts.samples, flush=True)
    ref, segs, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    a, b = _render_both(sc, cam, opts)
    assert np.array_equal(a.img, ref) and a.stats.segments == segs
    assert np.array_equal(b.img, ref) and b.stats.segments == segs


def test_triangle_mesh_full_frame_packets_equal_the_lane_traversal(gpu):
    """The benchmark frame (1920x1080, 64 spp, depth 8): packets and per-lane traversal agree bit for bit."""
    sc, cam, opts = presets.triangle_mesh()
    print("case packets triangle_mesh full frame", flush=True)
    a, b = _render_both(sc, cam, opts)
    assert np.array_equal(a.img, b.img)
    assert a.stats.segments == b.stats.segments
