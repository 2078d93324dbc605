"""Parity of the HIP path (through the C ABI) with the oracle, on the GPU.

The bar: the GPU computes the fp32 restatement bit for bit -- primary-hit ranks (the engine's depth-first
primitive order), hit distances, every pixel of every image and the number of closest-hit queries equal the
fp32 twin's exactly.  Against the fp64 engine precision the GPU's own primary hits and images are held to the
equal-seed contract of tests/precision.py (ranks, distances within 1e-5 relative, radiance within 1e-4
relative, every exception of a stated cause), with the per-workload bounds of tests/test_golden.py.  The
benchmark frames themselves (1920x1080 x 64 spp mesh, 1280x720 x 128 spp SDF + volume) are checked against the
fp32 twin on a strided subset of rows.
"""
import os

import numpy as np
import pytest

import oracle
from rayna_amd import Camera, Renderer, RenderOpts, RayError, presets
from rayna_amd import _abi as A

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SCENES = {
    "random_spheres": lambda: presets.random_spheres(width=64, height=36, samples=4),
    "mixed_primitives": lambda: presets.mixed_primitives(width=48, height=32, samples=4),
    "triangle_mesh_s4": lambda: presets.triangle_mesh(subdiv=4, width=48, height=27, samples=2),
    "sdf_volume": lambda: presets.sdf_volume(width=32, height=18, samples=2),
}


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_golden_bit_exact(gpu, name):
    sc, cam, opts = SCENES[name]()
    _print("case", name, opts)
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    with Renderer(sc, cam, opts) as r:
        assert r.tree_hash() == int(g["tree_hash"])
        rank, t = r.primary_hits(0)
        assert np.array_equal(rank, g["rank"])
        assert np.array_equal(t[rank != -1], g["t32"][rank != -1])
        out = r.render(layout="rowmajor")
        assert np.array_equal(out.img, g["img32"])
        assert out.stats.segments == int(g["segs32"])
    # against the fp64 engine precision, equal seeds: the GPU's own output held to the contract
    from tests import precision as P
    from tests.test_golden import GRAZING_MAX, RADIANCE_OFF_MAX
    tw = oracle.TwinScene(sc)
    prim = P.primary_check(tw, cam, opts, rank, t, g["rank64"], g["t64"])
    rad = P.radiance_check(tw, sc, cam, opts, out.img, g["img64"])
    _print(name, "vs fp64:", prim, rad)
    assert prim["rank_same"] >= P.RANK_MIN
    assert prim["t_grazing"] <= GRAZING_MAX * prim["hits"]
    assert rad["off"] <= RADIANCE_OFF_MAX[name] * rad["pixels"]


def test_engine_sphere_colours_on_gpu(gpu):
    albedo = (0.3, 0.55, 0.8)
    for z, want in ((0.0, "black"), (-1.01, "albedo"), (1.01, "white")):
        sc, cam, opts = presets.sphere_colours(albedo, z, 64, 64)
        _print("case sphere_colours z", z)
        with Renderer(sc, cam, opts) as r:
            img = r.render().img
        if want == "black":
            assert np.all(img == 0)
        elif want == "white":
            assert np.all(img == 1)
        else:
            a = np.float32(albedo)
            acc = np.zeros(3, np.float32)
            for _ in range(10):
                acc = acc + a
            assert np.all(img == acc / np.float32(10))
    # the engine test's fractions at its own options (simple_renders.rs:129-141)
    sc, cam, opts = presets.sphere_colours(albedo, -2.7)
    with Renderer(sc, cam, opts) as r:
        px = r.render().img.reshape(-1, 3)
    close = lambda t: np.all(np.abs(px - np.float32(t)) < 0.1, axis=1)
    white, black = close((1, 1, 1)), close((0, 0, 0))
    colour = close(albedo) & ~white & ~black
    assert colour.mean() >= 0.6 and 0.2 <= white.mean() <= 0.4 and black.mean() == 0
    assert (~(white | black | colour)).mean() <= 0.05


def test_accumulation_and_setters(gpu):
    sc, cam, opts = presets.random_spheres(width=40, height=24, samples=3)
    tw = oracle.TwinScene(sc)
    f0 = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, frame=0)[0]
    f1 = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, frame=1)[0]
    with Renderer(sc, cam, opts) as r:
        a = r.render(layout="rowmajor")
        assert a.stats.accum_frames == 1 and np.array_equal(a.img, f0)
        b = r.render(layout="rowmajor")
        assert b.stats.accum_frames == 2
        # AccumulationValue: f32 running sum, f64 count, mean rounded to f32 (accum_buffer.rs:37-41)
        want = ((f0 + f1).astype(np.float64) / 2.0).astype(np.float32)
        assert np.array_equal(b.img, want)
        r.clear_accumulation()
        c = r.render(layout="rowmajor")
        assert c.stats.accum_frames == 1 and np.array_equal(c.img, f0)
        r.render()
        r.set_camera(cam)
        assert r.render(layout="rowmajor").stats.accum_frames == 1
        r.set_options(opts)
        d = r.render(layout="rowmajor")
        assert d.stats.accum_frames == 1 and np.array_equal(d.img, f0)
        r.set_scene(sc)
        assert np.array_equal(r.render(layout="rowmajor").img, f0)
        # layouts: the engine's Image is (width, height), indexed [x][y]
        r.clear_accumulation()
        xm = r.render(layout="xmajor").img
        assert xm.shape == (40, 24, 3) and np.array_equal(xm.transpose(1, 0, 2), f0)


def test_chunking_and_tuning_leave_the_image_unchanged(gpu):
    """The same frame traced as one chunk of paths or as several (and with other launch shapes) is bit for
    bit the same image, with the same closest-hit count, as the oracle's."""
    sc, cam, opts = presets.mixed_primitives(width=96, height=64, samples=12)
    ref, segs, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    with Renderer(sc, cam, opts) as r:
        for tune in (dict(), dict(chunk_slots=1 << 14), dict(chunk_slots=1 << 16, refill_lanes=64, wave_batch=1),
                     dict(refill_lanes=1, wave_batch=4096, shade_blocks_per_cu=1)):
            print("case chunking", tune, flush=True)
            r.set_tuning(**tune)
            r.clear_accumulation()
            out = r.render(layout="rowmajor")
            assert np.array_equal(out.img, ref), tune
            assert out.stats.segments == segs, tune
            if tune.get("chunk_slots"):
                assert out.stats.extend_launches > opts.ray_depth + 1


@pytest.mark.parametrize("spp", [1, 7, 64, 300, 4200])
def test_samples_per_pixel_and_resolve_staging(gpu, spp):
    """A pixel's samples are neighbouring slots and the resolve sums them in sample order: through LDS rows of
    odd stride (up to 4159 samples per pixel in a chunk) or straight from HBM beyond; single frames and a batch
    agree with the oracle bit for bit."""
    from dataclasses import replace

    from rayna_amd.batch import BatchRenderer, Request
    sc, cam, opts = presets.mixed_primitives(width=5, height=3, samples=spp)
    opts = replace(opts, ray_depth=4)
    print("case spp", spp, flush=True)
    ref, segs, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    with Renderer(sc, cam, opts) as r:
        out = r.render(layout="rowmajor")
        assert np.array_equal(out.img, ref) and out.stats.segments == segs
    small = replace(opts, width=4, height=2, samples=max(1, spp // 3), seed=5)
    ref2, _, _, _ = oracle.TwinScene(sc).render(cam, small, precision=32, box_mode=oracle.BOX_TIGHT)
    with BatchRenderer([sc]) as b:
        imgs, st = b.render([Request(0, cam, opts), Request(0, cam, small)])
    assert np.array_equal(imgs[0], ref) and np.array_equal(imgs[1], ref2)


def test_invalid_camera_renders_the_checkerboard(gpu):
    sc, cam, opts = presets.random_spheres(width=40, height=24, samples=1)
    bad = Camera(cam.pos, cam.fwd, 0.0, cam.focus_dist)
    with Renderer(sc, bad, opts) as r:
        img = r.render(layout="xmajor").img
    w, h = 5, 4  # ilog2(40), ilog2(24)
    assert img.shape == (w, h, 3)
    for x in range(w):
        for y in range(h):
            want = [0, 0, 0] if (x + y) % 2 == 0 else [1, 0, 1]
            assert list(img[x, y]) == want


def test_unsupported_and_invalid_options(gpu):
    sc, cam, opts = presets.random_spheres(width=8, height=8, samples=1)
    with pytest.raises(RayError) as e:
        Renderer(sc, cam, RenderOpts(8, 8, 1, 5, ray_branching=2))
    assert e.value.status == A.RAYNA_UNSUPPORTED
    with pytest.raises(RayError) as e:
        Renderer(sc, cam, RenderOpts(0, 8, 1, 5))
    assert e.value.status == A.RAYNA_INVALID_ARGUMENT


def test_triangle_mesh_full_size_primary_hits_and_one_spp(gpu):
    """The 1080p workload at full resolution: every primary hit of sample 0 and a whole 1-spp frame equal the
    fp32 twin's bit for bit."""
    sc, cam, opts = presets.triangle_mesh(samples=1)
    _print("case triangle_mesh 1920x1080 1 spp")
    tw = oracle.TwinScene(sc)
    with Renderer(sc, cam, opts) as r:
        rank, t = r.primary_hits(0)
        rk, d = tw.primary_hits(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
        assert np.array_equal(rank, rk)
        assert np.array_equal(t[rank != -1], d[rank != -1].astype(np.float32))
        out = r.render(layout="rowmajor")
    ref, segs, _, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, threads=16)
    assert np.array_equal(out.img, ref)
    assert out.stats.segments == segs


def test_triangle_mesh_full_frame_properties(gpu):
    """The benchmark frame (1920x1080, 64 spp, depth 8): finite non-negative radiance, one closest-hit query per
    path vertex, and a re-render after clear_accumulation reproduces the frame bit for bit."""
    sc, cam, opts = presets.triangle_mesh()
    _print("case triangle_mesh full frame")
    with Renderer(sc, cam, opts) as r:
        a = r.render(layout="rowmajor")
        r.clear_accumulation()
        b = r.render(layout="rowmajor")
    assert np.isfinite(a.img).all() and (a.img >= 0).all()
    n = 1920 * 1080 * 64
    assert a.stats.samples == n and n <= a.stats.segments <= 9 * n
    assert np.array_equal(a.img, b.img) and a.stats.segments == b.stats.segments
    # mean over pixels agrees with a 1-spp twin estimate on a pixel subset within Monte Carlo error
    assert a.img.mean() > 0


@pytest.mark.parametrize("workload,row_step", [("triangle_mesh", 54), ("sdf_volume", 80)])
def test_benchmark_frame_rows_equal_the_oracle(gpu, workload, row_step):
    """The benchmark frames at their full size and sample count (1920x1080 x 64 spp with camera-ray packets;
    1280x720 x 128 spp), every row_step-th row against the fp32 twin bit for bit."""
    sc, cam, opts = getattr(presets, workload)()
    _print("case benchmark frame rows", workload, opts.width, opts.height, opts.samples, "every", row_step)
    with Renderer(sc, cam, opts) as r:
        out = r.render(layout="rowmajor")
    ref, _, secs, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT,
                                                  row_start=row_step // 2, row_step=row_step, threads=16)
    _print(workload, "oracle rows", ref.shape[0], f"{secs:.1f} s")
    assert np.array_equal(out.img[row_step // 2::row_step], ref)
