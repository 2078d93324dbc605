"""The Renderer: rayna_engine's `Renderer` API (rayna_engine/src/render/renderer.rs:36-264) over librayna_amd.so.

    r = Renderer.new_from(scene, camera, options)      # Renderer::new_from (renderer.rs:86-107)
    render = r.render()                                 # Renderer::render (renderer.rs:224-264)
    render.img          # float32 (width, height, 3): the engine's Image shape, indexed [x][y]
    render.stats        # RenderStats: duration, accum_frames, samples, segments, kernel timings
    r.set_scene / set_camera / set_options / clear_accumulation

All per-ray work runs in the HIP kernels of the library; there is no CPU fallback.  A missing library or GPU
raises RayError.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
# RAYNA_AMD_LIB selects another build of the library (e.g. a tuning variant); the default is the in-tree build.
LIB_PATH = os.environ.get("RAYNA_AMD_LIB", os.path.join(_HERE, "librayna_amd.so"))
_lib: Optional[C.CDLL] = None

# every symbol declared in include/rayna_gpu.h
EXPORTS = ("rayna_create", "rayna_set_scene", "rayna_set_camera", "rayna_set_options", "rayna_clear_accumulation",
           "rayna_render", "rayna_render_srgb8", "rayna_set_material_colours", "rayna_adjoint", "rayna_set_tuning",
           "rayna_set_validate", "rayna_validate_counts", "rayna_adjoint_static", "rayna_batch_create",
           "rayna_batch_set_tuning", "rayna_render_batch", "rayna_batch_destroy", "rayna_batch_last_error",
           "rayna_primary_hits", "rayna_dfs_order", "rayna_tree_hash", "rayna_scene_prepare", "rayna_scene_tree", "rayna_destroy",
           "rayna_measure_valu_peak", "rayna_debug_sqrt_mismatches", "rayna_debug_fail_launches", "rayna_last_error")


# RAYNA_CHECK_* of include/rayna_gpu.h, in order: the invariants of the engine's shared/validate.rs
VALIDATE_CHECKS = ("ray", "hit_dist", "hit_pos", "hit_normal", "hit_uv", "colour", "sample", "pixel")


class RayError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{A.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """The renderer library (loaded once).  Fails loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rayna_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.rayna_create.restype = C.c_int
        L.rayna_create.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc),
                                   C.POINTER(vp)]
        L.rayna_set_scene.restype = C.c_int
        L.rayna_set_scene.argtypes = [vp, C.POINTER(A.SceneDesc)]
        L.rayna_set_camera.restype = C.c_int
        L.rayna_set_camera.argtypes = [vp, C.POINTER(A.CameraDesc)]
        L.rayna_set_options.restype = C.c_int
        L.rayna_set_options.argtypes = [vp, C.POINTER(A.RenderOptsDesc)]
        L.rayna_clear_accumulation.restype = C.c_int
        L.rayna_clear_accumulation.argtypes = [vp]
        L.rayna_render.restype = C.c_int
        L.rayna_render.argtypes = [vp, C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_uint32),
                                   C.POINTER(A.RenderStats)]
        L.rayna_render_srgb8.restype = C.c_int
        L.rayna_render_srgb8.argtypes = [vp, C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(A.RenderStats)]
        L.rayna_set_material_colours.restype = C.c_int
        L.rayna_set_material_colours.argtypes = [vp, C.POINTER(C.c_float), C.c_uint32]
        L.rayna_adjoint.restype = C.c_int
        L.rayna_adjoint.argtypes = [vp, C.c_uint32, C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_double),
                                    C.c_uint32, C.POINTER(A.RenderStats)]
        L.rayna_set_tuning.restype = C.c_int
        L.rayna_set_tuning.argtypes = [vp, C.POINTER(A.Tuning)]
        L.rayna_adjoint_static.restype = C.c_int
        L.rayna_adjoint_static.argtypes = [vp, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32]
        L.rayna_set_validate.restype = C.c_int
        L.rayna_set_validate.argtypes = [vp, C.c_int32]
        L.rayna_validate_counts.restype = C.c_int
        L.rayna_validate_counts.argtypes = [vp, C.POINTER(C.c_uint64), C.c_int32]
        L.rayna_batch_create.restype = C.c_int
        L.rayna_batch_create.argtypes = [C.POINTER(A.SceneDesc), C.c_int32, C.POINTER(vp)]
        L.rayna_batch_set_tuning.restype = C.c_int
        L.rayna_batch_set_tuning.argtypes = [vp, C.POINTER(A.Tuning)]
        L.rayna_render_batch.restype = C.c_int
        L.rayna_render_batch.argtypes = [vp, C.POINTER(A.Request), C.c_int32, C.POINTER(A.BatchStats)]
        L.rayna_batch_destroy.restype = None
        L.rayna_batch_destroy.argtypes = [vp]
        L.rayna_batch_last_error.restype = C.c_char_p
        L.rayna_batch_last_error.argtypes = [vp]
        L.rayna_primary_hits.restype = C.c_int
        L.rayna_primary_hits.argtypes = [vp, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.rayna_dfs_order.restype = C.c_int
        L.rayna_dfs_order.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                      C.POINTER(C.c_int32)]
        L.rayna_tree_hash.restype = C.c_int
        L.rayna_tree_hash.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.rayna_scene_prepare.restype = C.c_int
        L.rayna_scene_prepare.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32)]
        L.rayna_scene_tree.restype = C.c_int
        L.rayna_scene_tree.argtypes = [C.POINTER(A.SceneDesc), C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]
        L.rayna_destroy.restype = None
        L.rayna_destroy.argtypes = [vp]
        L.rayna_measure_valu_peak.restype = C.c_int
        L.rayna_measure_valu_peak.argtypes = [C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.rayna_debug_sqrt_mismatches.restype = C.c_int
        L.rayna_debug_sqrt_mismatches.argtypes = [C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]
        L.rayna_debug_fail_launches.restype = C.c_int32
        L.rayna_debug_fail_launches.argtypes = [C.c_int32]
        L.rayna_last_error.restype = C.c_char_p
        L.rayna_last_error.argtypes = [vp]
        _lib = L
    return _lib


def measure_valu_peak(iters: int = 2048):
    """(wave64 VALU instructions per second, seconds) of the library's issue-rate calibration kernel on the
    current GPU: eight independent fp32 multiply-add chains per lane, no packed or fused forms."""
    rate, secs = C.c_double(0.0), C.c_double(0.0)
    _check(lib().rayna_measure_valu_peak(int(iters), C.byref(rate), C.byref(secs)))
    return rate.value, secs.value


def debug_sqrt_mismatches(first: int = 0, count: int = 1 << 32) -> int:
    """The number of float bit patterns in [first, first + count) on which the SDF distances' square root differs
    from the correctly rounded sqrtf (NaNs compare equal), counted on the current GPU."""
    out = C.c_uint64(0)
    _check(lib().rayna_debug_sqrt_mismatches(int(first), int(count), C.byref(out)))
    return int(out.value)


def debug_fail_launches(countdown: int) -> int:
    """Fault injection for recovery tests: the launch check `countdown` checks from now (0: the next) of any
    renderer reports a launch failure (RayError with status DEVICE_ERROR), after which that renderer raises
    RENDERER_LOST until it is replaced (Renderer.recreated).  No GPU fault is made.  countdown < 0 switches it off.
    Returns the previous setting (-1 when off)."""
    return int(lib().rayna_debug_fail_launches(int(countdown)))


def _check(status: int, handle=None):
    if status != A.RAYNA_OK:
        msg = lib().rayna_last_error(handle)
        raise RayError(status, msg.decode() if msg else "")


@dataclass
class RenderStats:
    """render/render.rs:4-15, plus the GPU's own counters."""
    duration: float
    accum_frames: int
    samples: int
    segments: int
    kernel_ms: float
    extend_ms: float
    shade_ms: float
    extend_launches: int
    extend_camera_ms: float = 0.0  # of extend_ms: the camera rays' (depth 0) launches


@dataclass
class Render:
    img: np.ndarray
    stats: RenderStats


def scene_tree(scene):
    """Host-only: the flattened traversal tree (rayna_scene_tree) as numpy arrays: node words (n, 8), item words
    (n, 2), transformed-object words (n, 2), root, per-lane stack need, packet stack need."""
    d = scene.desc()
    counts = (C.c_int32 * 3)()
    root, need, pneed = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    _check(lib().rayna_scene_tree(C.byref(d), None, 0, None, 0, None, 0, counts, C.byref(root), C.byref(need),
                                  C.byref(pneed)))
    nodes = np.zeros((counts[0], 8), np.int32)
    items = np.zeros((counts[1], 2), np.int32)
    xforms = np.zeros((counts[2], 2), np.int32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    _check(lib().rayna_scene_tree(C.byref(d), p(nodes), len(nodes), p(items), len(items), p(xforms), len(xforms),
                                  counts, C.byref(root), C.byref(need), C.byref(pneed)))
    return nodes, items, xforms, int(root.value), int(need.value), int(pneed.value)


def prepare_scene(scene):
    """Host-only BVH construction (no GPU): (tree hash, depth-first (object, local) arrays, stack need)."""
    d = scene.desc()
    h = C.c_uint64(0)
    n = C.c_int32(0)
    need = C.c_int32(0)
    _check(lib().rayna_scene_prepare(C.byref(d), C.byref(h), None, None, 0, C.byref(n), C.byref(need)))
    obj = np.zeros(n.value, np.int32)
    loc = np.zeros(n.value, np.int32)
    _check(lib().rayna_scene_prepare(C.byref(d), C.byref(h), obj.ctypes.data_as(C.POINTER(C.c_int32)),
                                     loc.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n), C.byref(need)))
    return int(h.value), obj, loc, int(need.value)


class Renderer:
    def __init__(self, scene, camera, options, num_threads: int = 0):
        self._scene, self._camera, self._options = scene, camera, options
        self._h = C.c_void_p()
        self._keep = [scene.desc(), camera.desc(), options.desc()]
        _check(lib().rayna_create(C.byref(self._keep[0]), C.byref(self._keep[1]), C.byref(self._keep[2]),
                                  C.byref(self._h)))
        self.num_threads = num_threads

    @classmethod
    def new_from(cls, scene, camera, options, num_threads: int = 0):
        return cls(scene, camera, options, num_threads)

    def recreated(self) -> "Renderer":
        """A new renderer of this one's scene, camera and options, after this one is closed: the recovery from
        RENDERER_LOST, as the engine's UI restarts its render worker from the scene, camera and options it
        keeps (rayna_ui/src/app.rs:483-495).  Tuning and validation go back to their defaults."""
        self.close()
        return Renderer.new_from(self._scene, self._camera, self._options, self.num_threads)

    def close(self):
        if self._h:
            lib().rayna_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # getters (renderer.rs:46-51)
    @property
    def scene(self):
        return self._scene

    @property
    def camera(self):
        return self._camera

    @property
    def options(self):
        return self._options

    def set_scene(self, scene):
        d = scene.desc()
        _check(lib().rayna_set_scene(self._h, C.byref(d)), self._h)
        self._scene = scene
        self._keep[0] = d

    def set_camera(self, camera):
        d = camera.desc()
        _check(lib().rayna_set_camera(self._h, C.byref(d)), self._h)
        self._camera = camera

    def set_options(self, options):
        d = options.desc()
        _check(lib().rayna_set_options(self._h, C.byref(d)), self._h)
        self._options = options

    def set_tuning(self, chunk_slots: int = 0, refill_lanes: int = 0, wave_batch: int = 0,
                   shade_blocks_per_cu: int = 0, packet_depths: int = 0):
        """Launch shape of the wavefront kernels (0 keeps a default; packet_depths -1: no packets).  A new
        chunk_slots re-plans the path buffers and clears accumulation."""
        t = A.Tuning(chunk_slots, refill_lanes, wave_batch, shade_blocks_per_cu, 0, 0, 0, packet_depths)
        _check(lib().rayna_set_tuning(self._h, C.byref(t)), self._h)

    def set_validate(self, enable: bool = True):
        """Count the engine's debug invariants (shared/validate.rs) in later PBR frames; on clears the counts."""
        _check(lib().rayna_set_validate(self._h, 1 if enable else 0), self._h)

    def validate_counts(self, reset: bool = False) -> dict:
        """Violations per check (VALIDATE_CHECKS) since validation was turned on or last reset."""
        c = (C.c_uint64 * len(VALIDATE_CHECKS))()
        _check(lib().rayna_validate_counts(self._h, c, 1 if reset else 0), self._h)
        return dict(zip(VALIDATE_CHECKS, (int(v) for v in c)))

    def clear_accumulation(self):
        _check(lib().rayna_clear_accumulation(self._h), self._h)

    def set_thread_count(self, num_threads: int):
        """The GPU renderer has no host worker pool; the count is kept for API compatibility."""
        self.num_threads = num_threads

    def render(self, layout: str = "xmajor", to_host: bool = True) -> Render:
        o = self._options
        img = np.zeros((o.width * o.height * 3,), np.float32) if to_host else None
        dims = (C.c_uint32 * 2)()
        st = A.RenderStats()
        lay = A.LAYOUT_XMAJOR if layout == "xmajor" else A.LAYOUT_ROWMAJOR
        ptr = img.ctypes.data_as(C.POINTER(C.c_float)) if to_host else None
        _check(lib().rayna_render(self._h, ptr, lay, dims, C.byref(st)), self._h)
        w, h = int(dims[0]), int(dims[1])
        if img is not None:
            img = img[: w * h * 3]
            img = img.reshape(w, h, 3) if lay == A.LAYOUT_XMAJOR else img.reshape(h, w, 3)
        stats = RenderStats(st.duration_s, int(st.accum_frames), int(st.samples), int(st.segments), st.kernel_ms,
                            st.extend_ms, st.shade_ms, int(st.extend_launches), st.extend_camera_ms)
        return Render(img, stats)

    def render_srgb8(self) -> Render:
        """render() followed by the UI's 8-bit conversion (rayna_ui/src/ext/img_ext.rs:15-45) on the GPU:
        (height, width, 4) uint8 RGBA, gamma 1/2.2."""
        o = self._options
        img = np.zeros(o.width * o.height * 4, np.uint8)
        dims = (C.c_uint32 * 2)()
        st = A.RenderStats()
        _check(lib().rayna_render_srgb8(self._h, img.ctypes.data_as(C.POINTER(C.c_uint8)), dims, C.byref(st)),
               self._h)
        w, h = int(dims[0]), int(dims[1])
        stats = RenderStats(st.duration_s, int(st.accum_frames), int(st.samples), int(st.segments), st.kernel_ms,
                            st.extend_ms, st.shade_ms, int(st.extend_launches), st.extend_camera_ms)
        return Render(img[: w * h * 4].reshape(h, w, 4), stats)

    # ---- differentiation ----
    def set_material_colours(self, rgb):
        """Replace the material colours (albedo, or emission for lights), an (n, 3) array in material-table
        order (Scene.materials_list()).  Clears accumulation."""
        c = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        _check(lib().rayna_set_material_colours(self._h, c.ctypes.data_as(C.POINTER(C.c_float)), len(c)), self._h)

    def adjoint(self, d_image, frame: int = 0, layout: str = "xmajor"):
        """Path-replay gradient of a loss with respect to every material colour, (n_materials, 3) float64,
        given d_image = dloss/d(frame `frame`'s per-pixel mean) in `layout` ((width, height, 3) for "xmajor",
        (height, width, 3) for "rowmajor").  Returns (grad, RenderStats of the call)."""
        o = self._options
        d = np.ascontiguousarray(d_image, np.float32)
        if d.size != o.width * o.height * 3:
            raise ValueError(f"d_image has {d.size} values, expected {o.width * o.height * 3}")
        n = len(self._scene.materials_list())
        grad = np.zeros((n, 3), np.float64)
        st = A.RenderStats()
        lay = A.LAYOUT_XMAJOR if layout == "xmajor" else A.LAYOUT_ROWMAJOR
        _check(lib().rayna_adjoint(self._h, frame, d.ctypes.data_as(C.POINTER(C.c_float)), lay,
                                   grad.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(st)), self._h)
        stats = RenderStats(st.duration_s, int(st.accum_frames), int(st.samples), int(st.segments), st.kernel_ms,
                            st.extend_ms, st.shade_ms, int(st.extend_launches), st.extend_camera_ms)
        return grad, stats

    # ---- parity hooks ----
    def primary_hits(self, sample: int = 0):
        o = self._options
        rank = np.zeros(o.width * o.height, np.int32)
        t = np.zeros(o.width * o.height, np.float32)
        _check(lib().rayna_primary_hits(self._h, sample, rank.ctypes.data_as(C.POINTER(C.c_int32)),
                                        t.ctypes.data_as(C.POINTER(C.c_float))), self._h)
        return rank.reshape(o.height, o.width), t.reshape(o.height, o.width)

    def dfs_order(self):
        n = C.c_int32(0)
        _check(lib().rayna_dfs_order(self._h, None, None, 0, C.byref(n)), self._h)
        obj = np.zeros(n.value, np.int32)
        loc = np.zeros(n.value, np.int32)
        _check(lib().rayna_dfs_order(self._h, obj.ctypes.data_as(C.POINTER(C.c_int32)),
                                     loc.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n)), self._h)
        return obj, loc

    def tree_hash(self) -> int:
        h = C.c_uint64(0)
        _check(lib().rayna_tree_hash(self._h, C.byref(h)), self._h)
        return int(h.value)
