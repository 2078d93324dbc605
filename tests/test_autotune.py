"""The launch-shape autotuner's search (coordinate descent from the defaults, a change kept only when it beats
the best by more than the margin), driven by a stand-in batch whose device time is a known function of the
shape; and the request scaling it tunes on."""
from types import SimpleNamespace

from rayna_amd import _abi as A
from rayna_amd import autotune as AT
from rayna_amd.batch import Request
from rayna_amd.renderer import RayError
from rayna_amd.scene import Camera, RenderOpts


class _Batch:
    def __init__(self, cost):
        self.cost, self.cfg, self.calls = cost, None, 0

    def set_tuning(self, **cfg):
        if cfg.get("use_graph") and cfg.get("schedule") == "interleaved":  # as the batch refuses it
            raise RayError(A.RAYNA_UNSUPPORTED, "the interleaved schedule is not captured in a graph")
        self.cfg = cfg

    def render(self, requests, to_host=True):
        self.calls += 1
        return None, SimpleNamespace(kernel_ms=self.cost(self.cfg))


def test_descent_finds_the_separable_optimum():
    want = {"chunk_slots": 1 << 26, "refill_lanes": 48, "wave_batch": 192, "shade_blocks_per_cu": 8,
            "pipelines": 4, "use_graph": False, "schedule": "by_scene"}
    cost = lambda c: 100.0 + sum(10.0 for k, v in want.items() if c[k] != v)
    b = _Batch(cost)
    r = AT.tune(b, [], repeats=1)
    assert r.best == want and r.best_ms == 100.0 and r.default_ms == 100.0 + 10.0 * 6
    assert b.cfg == {k: v for k, v in want.items()}  # left on the best shape


def test_refused_shapes_are_skipped():
    # graph capture pays off, and the interleaved schedule more, but the two together are refused
    cost = lambda c: 100.0 - 5.0 * c["use_graph"] - 8.0 * (c["schedule"] == "interleaved")
    r = AT.tune(_Batch(cost), [], repeats=1)
    assert r.best["schedule"] == "interleaved" and not r.best["use_graph"] and r.best_ms == 92.0
    assert all(t["ms"] < float("inf") for t in r.trials)  # refused shapes are not trials


def test_noise_below_the_margin_keeps_the_defaults():
    b = _Batch(lambda c: 100.0 - 0.5 * (c["pipelines"] == 4) - 0.3 * (c["wave_batch"] == 64))
    r = AT.tune(b, [], repeats=1, margin=0.01)
    assert r.best == AT.DEFAULTS
    r = AT.tune(b, [], repeats=1, margin=0.0)
    assert r.best["pipelines"] == 4 and r.best["wave_batch"] == 64


def test_scaled_requests_keep_at_least_one_sample():
    cam = Camera(pos=(0.0, 0.0, 0.0), fwd=(0.0, 0.0, -1.0), v_fov=40.0, focus_dist=1.0)
    reqs = [Request(0, cam, RenderOpts(width=8, height=4, samples=s, ray_depth=3, seed=1)) for s in (1, 3, 64)]
    out = AT.scaled_requests(reqs, 4)
    assert [q.opts.samples for q in out] == [1, 1, 16]
    assert all(a.opts.width == b.opts.width and a.camera is b.camera for a, b in zip(reqs, out))
