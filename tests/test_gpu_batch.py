"""Batched rendering (rayna_render_batch) through the C ABI: every request's image is bit for bit the first
frame of a fresh Renderer with the same scene, camera and options (and so, through the single-frame parity
tests, the oracle's); requests spanning several chunks, chunks holding several requests, both layouts, an
invalid camera, graph capture and replay, and the shortest-first schedule."""
from dataclasses import replace

import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, presets
from rayna_amd.batch import BatchRenderer, Request
from rayna_amd.scene import Camera

pytestmark = pytest.mark.gpu


def _requests(n=10, seed=21):
    scenes, reqs = presets.batch_requests(seed=seed, n=n, mesh_subdiv=3, max_w=160, max_h=96, min_side=16,
                                          max_spp=12)
    reqs[1].layout = "xmajor"
    return scenes, reqs


def _single(scenes, q):
    with Renderer(scenes[q.scene], q.camera, q.opts) as r:
        out = r.render(layout=q.layout)
    return out.img, out.stats.segments


@pytest.mark.parametrize("tuning", [dict(), dict(chunk_slots=1 << 16), dict(chunk_slots=1 << 16, use_graph=True),
                                    dict(schedule="shortest_first", refill_lanes=8, wave_batch=32, shade_blocks_per_cu=4),
                                    dict(schedule="by_scene", chunk_slots=1 << 17),
                                    dict(chunk_slots=1 << 16, pipelines=1), dict(chunk_slots=1 << 16, pipelines=3),
                                    dict(chunk_slots=1 << 16, pipelines=4, use_graph=True),
                                    dict(chunk_slots=1 << 16, pipelines=3, schedule="interleaved"),
                                    dict(chunk_slots=1 << 16, pipelines=2, schedule="interleaved")])
def test_batch_matches_single_renders(gpu, tuning):
    scenes, reqs = _requests()
    print("case batch", tuning, flush=True)
    with BatchRenderer(scenes) as b:
        b.set_tuning(**tuning)
        imgs, st = b.render(reqs)
        if tuning.get("use_graph"):
            for k in range(3):  # replays of the captured graph, with eager work of the program in between
                imgs2, _ = b.render(reqs)
                for a, c in zip(imgs, imgs2):
                    assert np.array_equal(a, c), f"replay {k}"
                _single(scenes, reqs[k])
    segs = 0
    for i, (q, img) in enumerate(zip(reqs, imgs)):
        ref, s = _single(scenes, q)
        segs += s
        assert img.shape == ref.shape, i
        assert np.array_equal(img, ref), f"request {i} differs from its single render"
    assert st.segments == segs
    assert st.samples == sum(q.opts.width * q.opts.height * q.opts.samples for q in reqs)
    if tuning.get("chunk_slots"):
        assert st.chunks > 1


def test_interleaved_schedule_is_not_captured(gpu):
    """Graph capture of the interleaved schedule is refused before anything is traced (a captured interleaved
    batch once ended its process with a segmentation fault in the runtime)."""
    from rayna_amd import RayError
    from rayna_amd import _abi as A
    scenes, _ = _requests(n=2)
    with BatchRenderer(scenes) as b:
        with pytest.raises(RayError) as e:
            b.set_tuning(schedule="interleaved", use_graph=True)
        assert e.value.status == A.RAYNA_UNSUPPORTED


def test_batch_against_oracle_and_invalid_camera(gpu):
    scenes, reqs = _requests(n=4, seed=5)
    bad = Request(0, Camera(pos=(0, 0, 0), fwd=(0, 0, 0), v_fov=40.0, focus_dist=1.0), reqs[0].opts)
    with BatchRenderer(scenes) as b:
        imgs, st = b.render(reqs + [bad])
    for q, img in zip(reqs, imgs):
        tw = oracle.TwinScene(scenes[q.scene])
        ref, _, _, _ = tw.render(q.camera, q.opts, precision=32, box_mode=oracle.BOX_TIGHT)
        if q.layout == "xmajor":
            ref = ref.transpose(1, 0, 2)
        assert np.array_equal(img, ref)
    with Renderer(scenes[0], bad.camera, bad.opts) as r:
        chk = r.render(layout="rowmajor").img
    w, h = chk.shape[1], chk.shape[0]
    assert np.array_equal(imgs[-1].reshape(-1)[: w * h * 3], chk.reshape(-1))


def test_batch_latency_order(gpu):
    """With shortest-first scheduling the smallest request finishes first."""
    scenes, reqs = _requests(n=12, seed=9)
    with BatchRenderer(scenes) as b:
        b.set_tuning(chunk_slots=1 << 16, schedule="shortest_first")
        _, st = b.render(reqs)
    work = np.array([q.opts.width * q.opts.height * q.opts.samples for q in reqs])
    assert st.request_ms[np.argmin(work)] <= st.request_ms.min() + 1e-9
    assert (st.request_ms > 0).all() and st.request_ms.max() <= st.kernel_ms + 1e-6
