"""rayna_adjoint_static (the adjoint that allocates nothing: its gradient copies live in a static device
workspace) called through the C ABI with device buffers, against rayna_adjoint on the same renderer: two calls
in a row, the second on a smaller frame with fewer materials, so the shared workspace must be cleared between
them; the same bar as the adjoint's parity tests (float64 atomics: normwise 1e-9)."""
import ctypes as C

import numpy as np
import pytest

from rayna_amd import Renderer, presets
from rayna_amd import _abi as A
from rayna_amd.renderer import RayError, lib

pytestmark = pytest.mark.gpu


def _static_adjoint(r, d_rowmajor, n_materials, frame=0):
    import torch
    d_dev = torch.from_numpy(np.ascontiguousarray(d_rowmajor, np.float32)).to("cuda")
    g_dev = torch.full((n_materials * 3,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = lib().rayna_adjoint_static(r._h, frame, C.c_void_p(d_dev.data_ptr()), C.c_void_p(g_dev.data_ptr()),
                                    n_materials)
    if st != A.RAYNA_OK:
        raise RayError(st, lib().rayna_last_error(r._h).decode())
    torch.cuda.synchronize()
    return g_dev.cpu().numpy().reshape(n_materials, 3)


def test_adjoint_static_equals_adjoint_twice(gpu):
    cases = [presets.inverse_spheres(width=64, height=48, samples=4)[:3],
             presets.random_spheres(width=32, height=20, samples=2)]
    sizes = []
    for sc, cam, opts in cases:
        print("case adjoint_static", opts.width, opts.height, opts.samples, flush=True)
        n = len(sc.materials_list())
        d = np.random.default_rng(opts.width).standard_normal((opts.height, opts.width, 3)).astype(np.float32)
        with Renderer(sc, cam, opts) as r:
            r.render()
            g_ref, _ = r.adjoint(d, frame=0, layout="rowmajor")
            g = _static_adjoint(r, d, n)
        rel = float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
        assert rel <= 1e-9, rel
        sizes.append((opts.width * opts.height, n))
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]


def test_adjoint_static_rejects_what_it_cannot_do(gpu):
    """A wrong material count, and a scene with unbounded objects before any frame has set up the renderer's
    march buffers (the entry point would have to allocate them), are RAYNA_INVALID_ARGUMENT; after a frame the
    same call works."""
    sc, cam, opts = presets.sdf_volume(width=16, height=8, samples=1)
    n = len(sc.materials_list())
    d = np.random.default_rng(3).standard_normal((opts.height, opts.width, 3)).astype(np.float32)
    with Renderer(sc, cam, opts) as r:
        for m in (n + 1, n):
            with pytest.raises(RayError) as e:
                _static_adjoint(r, d, m)
            assert e.value.status == A.RAYNA_INVALID_ARGUMENT
        r.render()
        g_ref, _ = r.adjoint(d, frame=0, layout="rowmajor")
        g = _static_adjoint(r, d, n)
    assert np.linalg.norm(g - g_ref) <= 1e-9 * max(np.linalg.norm(g_ref), 1e-30)
