"""Autotuner for the wavefront launch shape of batched rendering.

The tunables are those of rayna_tuning (include/rayna_gpu.h): paths per chunk (the queue size), the extend
kernel's refill threshold (idle lanes before a wave takes new rays: the regeneration threshold) and its claim
size, shade blocks per compute unit (the block shape of the shade grid), the chunks in flight at once, hipGraph
capture and the schedule (by scene, or interleaved; a shape the batch refuses, such as a captured interleaved
schedule, is skipped).  The search is coordinate descent from the defaults: each tunable in turn is swept with the
others held, and the fastest value (median device time over `repeats` runs of the tuning batch) is kept if it
beats the current best by more than `margin` (a fraction; run-to-run noise should not move the shape).  Results
are returned, not stored.
"""
from __future__ import annotations

import statistics
from dataclasses import dataclass, field, replace
from typing import Dict, List, Sequence

from .batch import BatchRenderer, Request
from .renderer import RayError

SPACE: Dict[str, Sequence] = {
    "chunk_slots": (1 << 26, 1 << 27, 1 << 28, 1 << 29),
    "refill_lanes": (8, 16, 32, 48),
    "wave_batch": (64, 128, 192, 256),
    "shade_blocks_per_cu": (8, 16, 32),
    "pipelines": (1, 2, 3, 4),
    "use_graph": (False, True),
    "schedule": ("by_scene", "interleaved"),
}
DEFAULTS = {"chunk_slots": 1 << 29, "refill_lanes": 32, "wave_batch": 128, "shade_blocks_per_cu": 16,
            "pipelines": 3, "use_graph": False, "schedule": "interleaved"}


@dataclass
class TuneResult:
    best: Dict[str, object]
    best_ms: float
    default_ms: float
    trials: List[dict] = field(default_factory=list)


def scaled_requests(requests: List[Request], spp_divisor: int) -> List[Request]:
    """The same requests with samples divided by spp_divisor (at least 1): a cheaper batch of the same shape."""
    return [replace(q, opts=replace(q.opts, samples=max(1, q.opts.samples // spp_divisor))) for q in requests]


def tune(batch: BatchRenderer, requests: List[Request], repeats: int = 2, space: Dict[str, Sequence] = SPACE,
         margin: float = 0.01) -> TuneResult:
    cache: Dict[tuple, float] = {}

    def measure(cfg) -> float:
        key = tuple(sorted(cfg.items()))
        if key in cache:
            return cache[key]
        try:
            batch.set_tuning(**cfg)
        except RayError:  # a shape the batch refuses
            cache[key] = float("inf")
            return cache[key]
        times = []
        for _ in range(repeats + (1 if cfg["use_graph"] else 0)):  # a graph's first run includes its capture
            _, st = batch.render(requests, to_host=False)
            times.append(st.kernel_ms)
        if cfg["use_graph"]:
            times = times[1:]
        cache[key] = statistics.median(times)
        res.trials.append({**cfg, "ms": cache[key]})
        return cache[key]

    res = TuneResult(dict(DEFAULTS), 0.0, 0.0)
    best = dict(DEFAULTS)
    best_ms = measure(best)
    res.default_ms = best_ms
    for name, values in space.items():
        for v in values:
            cfg = {**best, name: v}
            ms = measure(cfg)
            if ms < best_ms * (1.0 - margin):
                best, best_ms = cfg, ms
    res.best, res.best_ms = best, best_ms
    batch.set_tuning(**best)
    return res
