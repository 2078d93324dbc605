"""The oracle: a CPU restatement ("twin") of rayna_engine's render hot path, loaded with ctypes.

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use it, as the
checker; the product package (rayna_amd) never imports it.  The C++ sources (twin*.hpp, twin_capi.cpp) cite the
engine files they restate; `make -C oracle` builds liboracle.so.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from rayna_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib: Optional[C.CDLL] = None

BOX_REFERENCE, BOX_TIGHT = 0, 1
PATH_SKY, PATH_ABSORBED, PATH_DEPTH = -1000, -1001, -1002
# the invariants of the engine's shared/validate.rs, in the order of twin::CHECK_* and RAYNA_CHECK_*
CHECKS = ("ray", "hit_dist", "hit_pos", "hit_normal", "hit_uv", "colour", "sample", "pixel")
INTEGRATOR_RECURSIVE, INTEGRATOR_FORWARD = 0, 1


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.twin_scene_create.restype = vp
        L.twin_scene_create.argtypes = [C.POINTER(A.SceneDesc), C.c_int]
        L.twin_scene_destroy.argtypes = [vp]
        L.twin_scene_tree_hash.restype = C.c_uint64
        L.twin_scene_tree_hash.argtypes = [vp]
        L.twin_scene_dfs_order.restype = C.c_int32
        L.twin_scene_dfs_order.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32]
        L.twin_scene_tree_walk.restype = C.c_int32
        L.twin_scene_tree_walk.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32]
        L.twin_scene_set_colours.argtypes = [vp, C.POINTER(C.c_float), C.c_int32]
        L.twin_to_srgb8.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint8)]
        L.twin_adjoint.restype = C.c_uint64
        L.twin_adjoint.argtypes = [vp, C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_int32, C.c_int32,
                                   C.c_uint32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]
        L.twin_camera_invalid.restype = C.c_int32
        L.twin_camera_invalid.argtypes = [C.POINTER(A.CameraDesc)]
        L.twin_render.restype = C.c_uint64
        L.twin_render.argtypes = [vp, C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_int32, C.c_int32,
                                  C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_float),
                                  C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.twin_primary_hits.argtypes = [vp, C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_int32,
                                        C.c_int32, C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.twin_trace_ray.restype = C.c_int32
        L.twin_trace_ray.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.twin_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.twin_math_f32.argtypes = [C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.c_int64]
        L.twin_sampler_f64.argtypes = [C.c_int32, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double)]
        L.twin_scatter_f64.restype = C.c_int32
        L.twin_scatter_f64.argtypes = [C.POINTER(A.Material), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.twin_sky.argtypes = [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_float)]
        L.twin_texture_value.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_float)]
        L.twin_scene_sky.argtypes = [vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_float)]
        L.twin_noise_permutation.argtypes = [C.c_uint32, C.POINTER(C.c_int32)]
        L.twin_camera_ray.argtypes = [C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.POINTER(C.c_double)]
        L.twin_sdf_eval.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_int64,
                                    C.POINTER(C.c_double)]
        L.twin_render_checked.restype = C.c_uint64
        L.twin_render_checked.argtypes = [vp, C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_int32,
                                          C.c_int32, C.c_uint32, C.c_int32, C.POINTER(C.c_float),
                                          C.POINTER(C.c_uint64)]
        L.twin_path_ranks.argtypes = [vp, C.POINTER(A.CameraDesc), C.POINTER(A.RenderOptsDesc), C.c_int32, C.c_int32,
                                      C.c_uint32, C.c_uint32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def to_srgb8(img_xmajor: np.ndarray) -> np.ndarray:
    """The UI's 8-bit conversion of an x-major (width, height, 3) image: (height, width, 4) uint8 RGBA."""
    img = np.ascontiguousarray(img_xmajor, np.float32)
    w, h = img.shape[0], img.shape[1]
    out = np.zeros((h, w, 4), np.uint8)
    lib().twin_to_srgb8(_p(img, C.c_float), w, h, _p(out, C.c_uint8))
    return out


class TwinScene:
    """The engine's object graph for a Scene, built by the twin (builder 'sweep' or 'literal')."""

    def __init__(self, scene, builder: str = "sweep"):
        self.scene = scene
        self._desc = scene.desc()
        self.h = lib().twin_scene_create(C.byref(self._desc), 1 if builder == "literal" else 0)

    def __del__(self):
        if getattr(self, "h", None):
            lib().twin_scene_destroy(self.h)
            self.h = None

    def tree_hash(self) -> int:
        return int(lib().twin_scene_tree_hash(self.h))

    def dfs_order(self):
        n = lib().twin_scene_dfs_order(self.h, None, None, 0)
        o = np.zeros(n, np.int32)
        loc = np.zeros(n, np.int32)
        lib().twin_scene_dfs_order(self.h, _p(o, C.c_int32), _p(loc, C.c_int32), n)
        return o, loc

    def tree_walk(self, mesh: int = -1):
        n = lib().twin_scene_tree_walk(self.h, mesh, None, None, 0)
        items = np.zeros(n, np.int32)
        nch = np.zeros(n, np.int32)
        lib().twin_scene_tree_walk(self.h, mesh, _p(items, C.c_int32), _p(nch, C.c_int32), n)
        return items, nch

    def render(self, cam, opts, precision: int = 64, box_mode: int = BOX_REFERENCE, integrator: Optional[int] = None,
               frame: int = 0, row_start: int = 0, row_step: int = 1, threads: int = 0, col_step: int = 1):
        """Per-pixel sample means (rows x cols x 3, float32), closest-hit query count, seconds, prim tests."""
        if integrator is None:
            integrator = INTEGRATOR_FORWARD if precision == 32 else INTEGRATOR_RECURSIVE
        n_rows = (opts.height - row_start + row_step - 1) // row_step
        n_cols = (opts.width + col_step - 1) // col_step
        out = np.zeros((n_rows, n_cols, 3), np.float32)
        secs = C.c_double(0)
        tests = C.c_uint64(0)
        cd, od = cam.desc(), opts.desc()
        segs = lib().twin_render(self.h, C.byref(cd), C.byref(od), precision, box_mode, integrator, frame, row_start,
                                 row_step, col_step, threads, _p(out, C.c_float), C.byref(secs), C.byref(tests))
        return out, int(segs), secs.value, int(tests.value)

    def render_checked(self, cam, opts, precision: int = 32, box_mode: Optional[int] = None, frame: int = 0,
                       threads: int = 0):
        """A frame by the forward integrator with the engine's debug invariants (shared/validate.rs) counted:
        (image rows x cols x 3, counts per CHECKS entry)."""
        if box_mode is None:
            box_mode = BOX_TIGHT if precision == 32 else BOX_REFERENCE
        out = np.zeros((opts.height, opts.width, 3), np.float32)
        checks = np.zeros(len(CHECKS), np.uint64)
        cd, od = cam.desc(), opts.desc()
        lib().twin_render_checked(self.h, C.byref(cd), C.byref(od), precision, box_mode, frame, threads,
                                  _p(out, C.c_float), _p(checks, C.c_uint64))
        return out, dict(zip(CHECKS, (int(v) for v in checks)))

    def set_colours(self, rgb):
        """Replace the material colours (albedo, or emission for lights): an (n_materials, 3) array."""
        c = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        lib().twin_scene_set_colours(self.h, _p(c, C.c_float), len(c))

    def adjoint(self, cam, opts, d_image, precision: int = 64, box_mode: int = BOX_TIGHT, frame: int = 0,
                threads: int = 0):
        """Path-replay gradient of a loss w.r.t. every material colour, (n_materials, 3) float64, given
        d_image = dloss/d(frame mean), (height, width, 3) row-major.  Returns (grad, replay queries, seconds)."""
        d = np.ascontiguousarray(d_image, np.float32).reshape(opts.height, opts.width, 3)
        grad = np.zeros((len(self.scene.materials_list()), 3), np.float64)
        secs = C.c_double(0)
        cd, od = cam.desc(), opts.desc()
        segs = lib().twin_adjoint(self.h, C.byref(cd), C.byref(od), precision, box_mode, frame, threads,
                                  _p(d, C.c_float), _p(grad, C.c_double), C.byref(secs))
        return grad, int(segs), secs.value

    def primary_hits(self, cam, opts, precision: int = 32, box_mode: int = BOX_TIGHT, sample: int = 0):
        n = opts.width * opts.height
        rank = np.zeros(n, np.int32)
        dist = np.zeros(n, np.float64)
        cd, od = cam.desc(), opts.desc()
        lib().twin_primary_hits(self.h, C.byref(cd), C.byref(od), precision, box_mode, sample, _p(rank, C.c_int32),
                                _p(dist, C.c_double))
        return rank.reshape(opts.height, opts.width), dist.reshape(opts.height, opts.width)

    def trace(self, origin, direction, precision: int = 64, box_mode: int = BOX_REFERENCE):
        o = np.ascontiguousarray(origin, np.float64)
        d = np.ascontiguousarray(direction, np.float64)
        out = np.zeros(11)
        mat = C.c_int32(-1)
        rank = lib().twin_trace_ray(self.h, precision, box_mode, _p(o, C.c_double), _p(d, C.c_double),
                                    _p(out, C.c_double), C.byref(mat))
        if rank == -1:
            return None
        return dict(rank=int(rank), t=out[0], pos=out[1:4], normal=out[4:7], ray_normal=out[7:10],
                    front_face=bool(out[10]), material=int(mat.value))

    def texture_value(self, tex: int, pos, pos_l=(0.0, 0.0, 0.0), uv=(0.0, 0.0), precision: int = 64):
        """The value of texture `tex` (0-based, the scene description's order) at a hit."""
        p, pl, u = (np.ascontiguousarray(v, np.float64) for v in (pos, pos_l, uv))
        out = np.zeros(3, np.float32)
        lib().twin_texture_value(self.h, tex, precision, _p(p, C.c_double), _p(pl, C.c_double), _p(u, C.c_double),
                                 _p(out, C.c_float))
        return out

    def sky(self, d, precision: int = 64):
        dd = np.ascontiguousarray(d, np.float64)
        out = np.zeros(3, np.float32)
        lib().twin_scene_sky(self.h, precision, _p(dd, C.c_double), _p(out, C.c_float))
        return out

    def path_ranks(self, cam, opts, x: int, y: int, precision: int = 64, box_mode: Optional[int] = None):
        """Per sample of pixel (x, y), frame 0: the tuple of closest-hit ranks along its path, ended by how the
        path ended (PATH_SKY, PATH_ABSORBED or PATH_DEPTH)."""
        if box_mode is None:
            box_mode = BOX_REFERENCE if precision == 64 else BOX_TIGHT
        cap = opts.ray_depth + 3
        codes = np.zeros((opts.samples, cap), np.int32)
        lens = np.zeros(opts.samples, np.int32)
        cd, od = cam.desc(), opts.desc()
        lib().twin_path_ranks(self.h, C.byref(cd), C.byref(od), precision, box_mode, x, y, _p(codes, C.c_int32), cap,
                              _p(lens, C.c_int32))
        return [tuple(int(v) for v in codes[k, :lens[k]]) for k in range(opts.samples)]

    def sdf_eval(self, sdf: int, points, precision: int = 64):
        p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        out = np.zeros(len(p))
        lib().twin_sdf_eval(self.h, sdf, precision, _p(p, C.c_double), len(p), _p(out, C.c_double))
        return out


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().twin_philox4x32(c, k, o)
    return list(o)


MATH_FNS = {"sin_turns": 0, "cos_turns": 1, "log": 2, "exp": 3, "cbrt01": 4, "sin": 5, "cos": 6, "atan2": 7,
            "acos": 8}


def math_f32(fn: str, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(np.zeros_like(x) if y is None else y, np.float32)
    out = np.zeros_like(x)
    lib().twin_math_f32(MATH_FNS[fn], _p(x, C.c_float), _p(y, C.c_float), _p(out, C.c_float), x.size)
    return out


def sampler_f64(kind: str, u):
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    out = np.zeros_like(u)
    lib().twin_sampler_f64({"sphere": 0, "ball": 1, "disk": 2}[kind], _p(u, C.c_double), len(u), _p(out, C.c_double))
    return out


def scatter_f64(material, ray, hit, u):
    m = material._desc()
    r = np.ascontiguousarray(ray, np.float64)
    h = np.ascontiguousarray(hit, np.float64)
    uu = np.ascontiguousarray(u, np.float64)
    out = np.zeros(6)
    ok = lib().twin_scatter_f64(C.byref(m), _p(r, C.c_double), _p(h, C.c_double), _p(uu, C.c_double),
                                _p(out, C.c_double))
    return bool(ok), out[:3], out[3:]


def noise_permutation(seed: int) -> np.ndarray:
    out = np.zeros(512, np.int32)
    lib().twin_noise_permutation(seed, _p(out, C.c_int32))
    return out


def sky(kind: int, d):
    dd = np.ascontiguousarray(d, np.float64)
    out = np.zeros(3, np.float32)
    lib().twin_sky(kind, _p(dd, C.c_double), _p(out, C.c_float))
    return out


def camera_ray(cam, opts, x, y, k):
    out = np.zeros(6)
    cd, od = cam.desc(), opts.desc()
    lib().twin_camera_ray(C.byref(cd), C.byref(od), x, y, k, _p(out, C.c_double))
    return out[:3], out[3:]


def camera_invalid(cam) -> bool:
    cd = cam.desc()
    return bool(lib().twin_camera_invalid(C.byref(cd)))
