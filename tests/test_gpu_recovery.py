"""Recovery from a failed launch, forced with rayna_debug_fail_launches (no GPU fault is made).

The engine's render cannot fail (rayna_engine/src/render/renderer.rs:224-264); its caller notices a dead render
worker (rayna_ui/src/integration/mod.rs:100-125) and builds a new one from the scene, camera and options it
keeps (rayna_ui/src/app.rs:483-495).  Here the failed call reports RAYNA_DEVICE_ERROR, every later call on that
renderer RAYNA_RENDERER_LOST, and Renderer.recreated() is the caller's restart: its frames are bit-identical to
a fresh renderer's and to the oracle's."""
import numpy as np
import pytest

import oracle
from rayna_amd import Renderer, RayError, presets
from rayna_amd import _abi as A
from rayna_amd.renderer import debug_fail_launches

pytestmark = pytest.mark.gpu


@pytest.fixture
def injection():
    debug_fail_launches(-1)
    yield
    debug_fail_launches(-1)


def _lost_calls(r, cam, opts):
    return {"render": lambda: r.render(), "render_srgb8": lambda: r.render_srgb8(),
            "set_camera": lambda: r.set_camera(cam), "set_options": lambda: r.set_options(opts),
            "clear_accumulation": lambda: r.clear_accumulation(), "validate_counts": lambda: r.validate_counts(),
            "primary_hits": lambda: r.primary_hits(), "adjoint": lambda: r.adjoint(np.zeros((opts.width, opts.height, 3)))}


def test_failed_render_loses_the_renderer_and_a_new_one_renders_the_same(gpu, injection):
    sc, cam, opts = presets.random_spheres(width=32, height=18, samples=4)
    print("case recovery render", flush=True)
    ref, _, _, _ = oracle.TwinScene(sc).render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT)
    r = Renderer(sc, cam, opts)
    other = Renderer(sc, cam, opts)
    try:
        first = r.render(layout="rowmajor").img
        assert np.array_equal(first, ref)
        assert debug_fail_launches(1) == -1  # the second launch check from now fails
        r.clear_accumulation()
        assert np.array_equal(r.render(layout="rowmajor").img, ref)  # check 1 passes
        with pytest.raises(RayError) as e:
            r.render()
        assert e.value.status == A.RAYNA_DEVICE_ERROR and "(injected)" in str(e.value), str(e.value)
        assert debug_fail_launches(-1) == -1  # spent
        for name, call in _lost_calls(r, cam, opts).items():
            with pytest.raises(RayError) as e2:
                call()
            assert e2.value.status == A.RAYNA_RENDERER_LOST, (name, str(e2.value))
            assert "(injected)" in str(e2.value) and "create a new one" in str(e2.value)
        # another renderer in the process is untouched
        assert np.array_equal(other.render(layout="rowmajor").img, ref)
        r = r.recreated()
        assert np.array_equal(r.render(layout="rowmajor").img, ref)
        assert r.render().stats.accum_frames == 2
    finally:
        r.close()
        other.close()


def test_failed_adjoint_and_srgb8_lose_the_renderer(gpu, injection):
    sc, cam, opts, _ = presets.inverse_spheres(width=24, height=24, samples=2)
    print("case recovery adjoint", flush=True)
    n = len(sc.materials_list())
    with Renderer(sc, cam, opts) as fresh:
        fresh.render()
        g_ref, _ = fresh.adjoint(np.ones((opts.width, opts.height, 3), np.float32))
    with Renderer(sc, cam, opts) as fresh:
        rgba_ref = fresh.render_srgb8().img
    r = Renderer(sc, cam, opts)
    try:
        r.render()
        debug_fail_launches(0)
        with pytest.raises(RayError) as e:
            r.adjoint(np.ones((opts.width, opts.height, 3), np.float32))
        assert e.value.status == A.RAYNA_DEVICE_ERROR and "adjoint" in str(e.value)
        r = r.recreated()
        r.render()
        g, _ = r.adjoint(np.ones((opts.width, opts.height, 3), np.float32))
        # the gradient copies are summed with atomics, in no fixed order
        assert g.shape == (n, 3) and np.linalg.norm(g - g_ref) <= 1e-9 * max(np.linalg.norm(g_ref), 1e-30)
        # rayna_render_srgb8 makes its render's launch check and one of its own: fail the second
        debug_fail_launches(1)
        with pytest.raises(RayError) as e:
            r.render_srgb8()
        assert e.value.status == A.RAYNA_DEVICE_ERROR and "render_srgb8" in str(e.value)
        with pytest.raises(RayError) as e:
            r.render()
        assert e.value.status == A.RAYNA_RENDERER_LOST
        r = r.recreated()
        assert np.array_equal(r.render_srgb8().img, rgba_ref)
    finally:
        r.close()
