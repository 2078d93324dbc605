"""Batched rendering: many requests (scene, camera, resolution, samples, seed) traced together on the GPU.

    b = BatchRenderer([scene_a, scene_b])
    images, stats = b.render([Request(0, cam, opts), Request(1, cam2, opts2), ...])

Each image is bit for bit what a fresh Renderer (render/renderer.rs:86-107) renders as its first frame with the
same scene, camera and options; the batch packs all requests' samples into chunks of paths and runs one
wavefront pipeline per chunk (include/rayna_gpu.h, rayna_render_batch).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import _abi as A
from .renderer import RayError, lib


SCHEDULES = {"given": 0, "shortest_first": 1, "by_scene": 2, "interleaved": 3}


@dataclass
class Request:
    scene: int          # index into the BatchRenderer's scenes
    camera: object      # scene.Camera
    opts: object        # scene.RenderOpts
    layout: str = "rowmajor"


@dataclass
class BatchStats:
    duration: float      # host seconds of the call
    kernel_ms: float     # device time of the batch
    samples: int
    segments: int
    chunks: int
    launches: int
    request_ms: np.ndarray  # per request: device time from the batch start until its image was final (eager)


class BatchRenderer:
    def __init__(self, scenes: Sequence):
        self.scenes = list(scenes)
        self._descs = [s.desc() for s in self.scenes]
        arr = (A.SceneDesc * len(self._descs))(*self._descs)
        self._h = C.c_void_p()
        st = lib().rayna_batch_create(arr, len(self._descs), C.byref(self._h))
        if st != A.RAYNA_OK:
            raise RayError(st, lib().rayna_batch_last_error(None).decode())
        self.use_graph = False

    def _check(self, st):
        if st != A.RAYNA_OK:
            raise RayError(st, lib().rayna_batch_last_error(self._h).decode())

    def set_tuning(self, chunk_slots: int = 0, refill_lanes: int = 0, wave_batch: int = 0,
                   shade_blocks_per_cu: int = 0, use_graph: bool = False, schedule: str = "given",
                   pipelines: int = 0, packet_depths: int = 0):
        """Launch shape (0 keeps a default), graph capture, the order requests are traced in ("given",
        "shortest_first", "by_scene": scene groups by total samples, shortest first within a group, or
        "interleaved": by scene, with one-scene chunks dealt in turn from the groups) and the chunks in flight
        at once (`pipelines`, 1..4, each on its own stream; default 3)."""
        t = A.Tuning(chunk_slots, refill_lanes, wave_batch, shade_blocks_per_cu, 1 if use_graph else 0,
                     SCHEDULES[schedule], pipelines, packet_depths)
        self._check(lib().rayna_batch_set_tuning(self._h, C.byref(t)))
        self.use_graph = use_graph

    def render(self, requests: List[Request], to_host: bool = True):
        n = len(requests)
        imgs = []
        reqs = (A.Request * max(1, n))()
        for i, q in enumerate(requests):
            o = q.opts
            lay = A.LAYOUT_XMAJOR if q.layout == "xmajor" else A.LAYOUT_ROWMAJOR
            img = np.zeros(o.width * o.height * 3, np.float32) if to_host else None
            imgs.append(img)
            reqs[i] = A.Request(q.scene, lay, q.camera.desc(), o.desc(),
                                img.ctypes.data_as(C.POINTER(C.c_float)) if to_host else None)
        req_ms = np.zeros(max(1, n), np.float64)
        st = A.BatchStats()
        st.request_ms = req_ms.ctypes.data_as(C.POINTER(C.c_double))
        self._check(lib().rayna_render_batch(self._h, reqs, n, C.byref(st)))
        out = []
        for q, img in zip(requests, imgs):
            if img is None:
                out.append(None)
                continue
            o = q.opts
            out.append(img.reshape(o.height, o.width, 3) if q.layout != "xmajor" else img.reshape(o.width, o.height, 3))
        stats = BatchStats(st.duration_s, st.kernel_ms, int(st.samples), int(st.segments), int(st.chunks),
                           int(st.launches), req_ms[:n].copy())
        return out, stats

    def close(self):
        if self._h:
            lib().rayna_batch_destroy(self._h)
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
