"""Benchmark of the render hot path on one MI355X: frames of the 1080p triangle-mesh workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload triangle_mesh|random_spheres|sdf_volume|rttnw|cornell]

A step is one `Renderer::render()` frame: every sample of every pixel traced through the wavefront kernels, with
the scene, BVH and path buffers already resident in HBM (the image stays on the device).  Prints one JSON line:
Msamples/s (whole frame), Mrays/s (closest-hit queries), the roofline of the dominant kernel (extend, the
closest-hit traversal) and the CPU baseline (the oracle's double-precision restatement of the engine's
renderer, timed on this host on a bounded pixel sample).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBPS = 8000.0
# Algorithmic bytes of one closest-hit query in the extend kernel: queue entry 4 B + ray origin and direction
# 24 B read, hit distance, reference and flags 12 B written.
EXTEND_BYTES_PER_SEGMENT = 40
# Algorithmic bytes of the whole frame per segment (extend + shade) and per sample (generate + resolve):
# extend 40 B; shade reads queue 4, ray 24, hit 12, throughput 12, radiance 12 and writes ray 24,
# throughput 12, radiance 12, queue entry 4 -> 116 B; generate writes the 24 B camera ray per sample, resolve
# reads 12 B.
FRAME_BYTES_PER_SEGMENT = 40 + 116
FRAME_BYTES_PER_SAMPLE = 24 + 12
# Counter-derived figures of the extend stage on the 1080p triangle-mesh frame, from rocprofv3 passes over two
# frames of 263,040,107 closest-hit queries each (132,710,400 of them camera rays) at the commit that added them;
# the times they are divided by are measured live.  Traffic (both extend kernels): what leaves L2 for the
# fabric, FETCH_SIZE 4.1536e7 KiB read as 64-byte units (x2.0: a 256 MiB copy reads 131,087 KiB) plus WRITE_SIZE
# 1.4767e7 KiB (x1.0: the copy writes 262,143 KiB); it counts hits in the Infinity Cache as well as HBM, so it
# bounds the HBM bytes from above.  VALU (k_extend, depths 1-8): SQ_INSTS_VALU 3.3124e10 wave64 instructions over
# 2 x 130,329,707 queries; lanes active SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU) = 7.9352e11 / (64 x
# 3.3368e10); 0.878 instructions per CU per cycle (GRBM_GUI_ACTIVE 1.1794e9 over 8 XCDs, 256 CUs).
EXTEND_COUNTER_FRAMES = 2 * 263040107
EXTEND_FABRIC_BYTES_PER_SEGMENT = {"triangle_mesh": (4.1536e7 * 2.0 + 1.4767e7) * 1024 / EXTEND_COUNTER_FRAMES}
K_EXTEND_VALU_PER_QUERY = {"triangle_mesh": 3.3124e10 / (2 * 130329707)}
K_EXTEND_LANES_ACTIVE = {"triangle_mesh": 7.9352e11 / (64 * 3.3368e10)}
K_EXTEND_VALU_PER_CU_CYCLE = {"triangle_mesh": 3.3124e10 / (1.1794e9 / 8 * 256)}
# VALU issue peak: one wave64 instruction per CU per cycle -- the rate the calibration kernel
# (rayna_measure_valu_peak: independent fp32 multiply-add chains, no packed or fused forms) reaches per active
# cycle (0.977 by the same counters) -- at MI355X's published 256 CUs and 2.4 GHz.
MI355X_CUS, MI355X_GHZ = 256, 2.4
VALU_PEAK_G = MI355X_CUS * MI355X_GHZ * 1.0


_T0 = time.perf_counter()


def _progress(msg: str):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench [{time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def valu_roofline(workload: str, camera_rays: int, segs: int, ext_ms: float, cam_ms: float):
    """k_extend (depths 1-8) against the VALU issue peak: its wave64 VALU instructions per closest-hit query
    (counted by rocprofv3 on this workload) times this run's queries at those depths, over this run's k_extend
    time (the extend launches' HIP events less the camera rays')."""
    per_q = K_EXTEND_VALU_PER_QUERY.get(workload)
    t = (ext_ms - cam_ms) * 1e-3
    if not per_q or t <= 0:
        return None
    from rayna_amd.renderer import measure_valu_peak
    calib, _ = measure_valu_peak()
    achieved = per_q * (segs - camera_rays) / t / 1e9
    lanes = K_EXTEND_LANES_ACTIVE[workload]
    return {"kernel": "k_extend (depths 1-8)", "achieved": achieved, "peak": VALU_PEAK_G,
            "unit": "G wave64 VALU instructions/s", "frac": achieved / VALU_PEAK_G, "insts_per_query": per_q,
            "lanes_active": lanes, "useful_frac": achieved / VALU_PEAK_G * lanes,
            "insts_per_cu_cycle_profiled": K_EXTEND_VALU_PER_CU_CYCLE[workload],
            "k_extend_ms_per_frame": None, "calibration_g_per_s": calib / 1e9,
            "peak_basis": "1 wave64 VALU instruction per CU per cycle (the calibration kernel's rate per active "
                          "cycle) x 256 CUs x 2.4 GHz"}


class _NoGc:
    """Collects, then pauses Python's cyclic collector for a timed region: a collection walks the whole heap
    (torch's objects included) and can stall the host, and so the GPU, for tens of milliseconds."""

    def __enter__(self):
        gc.collect()
        gc.disable()

    def __exit__(self, *exc):
        gc.enable()


def workload(name: str):
    from rayna_amd import presets
    if name == "triangle_mesh":
        return presets.triangle_mesh()
    if name == "random_spheres":
        return presets.random_spheres()
    if name == "sdf_volume":
        return presets.sdf_volume()
    if name == "rttnw":  # the engine's RTTNW_DEMO: scattering sphere medium, instances, noise texture
        return presets.rttnw()
    if name == "cornell":  # the engine's CORNELL: rotated boxes (transforms), one light
        return presets.cornell()
    raise SystemExit(f"unknown workload {name}")


def host_cpus():
    """Threads for the CPU legs (the host's CPU share for this GPU: at most 16), and what the host shows."""
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(16, share), {"cpu_model": model, "cpus_visible": os.cpu_count(), "cpus_in_affinity": share}


def cpu_baseline(tw, cam, opts, target_seconds: float, threads: int, host: dict):
    """The oracle's fp64 restatement of the engine renderer (engine BVH boxes, recursive integrator) on a
    strided pixel sample of the same frame: the median of three runs, each about a fifth of `target_seconds`
    of work; beside it the same sample by the fp32 restatement (the GPU's arithmetic, tight boxes, forward
    integrator) and by the fp64 one with tight boxes."""
    import statistics

    import oracle
    per_run = target_seconds / 5.0
    # calibration pass: a tiny sample
    row_step, col_step = max(1, opts.height // 4), max(1, opts.width // 8)
    _, _, secs, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE, row_step=row_step,
                              col_step=col_step, threads=threads)
    n_px = ((opts.height + row_step - 1) // row_step) * ((opts.width + col_step - 1) // col_step)
    per_px = max(secs / n_px, 1e-7)
    for _ in range(2):  # a tiny sample overstates the cost per pixel (idle threads), so refine once
        want_px = max(1.0, per_run / per_px)
        scale = max(1.0, (opts.width * opts.height / want_px) ** 0.5)
        row_step = max(1, int(round(scale)))
        col_step = max(1, int(round(scale)))
        out, segs, secs, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE,
                                       row_step=row_step, col_step=col_step, threads=threads)
        n_px = out.shape[0] * out.shape[1]
        per_px = max(secs / n_px, 1e-7)
        if secs > 0.4 * per_run or scale == 1.0:
            break
    runs = [secs]
    for _ in range(2):
        runs.append(tw.render(cam, opts, precision=64, box_mode=oracle.BOX_REFERENCE, row_step=row_step,
                              col_step=col_step, threads=threads)[2])
    med = statistics.median(runs)
    samples = n_px * opts.samples
    # the same sample with refitted tight boxes (identical image, the engine's culling removed), and in fp32
    out2, _, secs2, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT, row_step=row_step,
                                  col_step=col_step, threads=threads)
    _, _, secs32, _ = tw.render(cam, opts, precision=32, box_mode=oracle.BOX_TIGHT, row_step=row_step,
                                col_step=col_step, threads=threads)
    return {
        "value": samples / med / 1e6,
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"fp64 oracle, engine BVH boxes, every {row_step}th row x every {col_step}th column at "
                  f"{opts.samples} spp ({n_px} pixels, {samples} samples, {segs} closest-hit queries): median of "
                  f"3 runs ({', '.join(f'{x:.2f}' for x in runs)} s) on {threads} threads",
        **host,
        "tight_boxes_value": samples / secs2 / 1e6,
        "tight_boxes_identical_image": bool((out == out2).all()),
        "fp32_oracle_tight_boxes_value": samples / secs32 / 1e6,
    }


def rmse_leg(r, tw, cam, opts, threads: int, ref_spp: int = 16384, grid: int = 16):
    """Converged-image error of the frame: RMSE over a grid x grid strided pixel sample against an fp64
    oracle render of the same pixels at ref_spp samples with another seed, for the GPU frame and for the
    fp64 oracle's own frame at the benchmark's samples (equal RMSE: the GPU estimator is the engine's)."""
    import numpy as np
    from dataclasses import replace
    import oracle

    img = r.render(layout="rowmajor").img  # a fresh first frame (the renderer was cleared by the caller)
    rs, cs = max(1, opts.height // grid), max(1, opts.width // grid)
    ref, _, secs, _ = tw.render(cam, replace(opts, samples=ref_spp, seed=opts.seed + 1), precision=64,
                                box_mode=oracle.BOX_TIGHT, row_step=rs, col_step=cs, threads=threads)
    f64, _, _, _ = tw.render(cam, opts, precision=64, box_mode=oracle.BOX_TIGHT, row_step=rs, col_step=cs,
                             threads=threads)
    gpu = img[::rs, ::cs][: ref.shape[0], : ref.shape[1]]
    rm = lambda a: float(np.sqrt(((a.astype(np.float64) - ref.astype(np.float64)) ** 2).mean()))
    return {"reference": f"fp64 oracle, {ref_spp} spp, seed {opts.seed + 1}, every {rs}th row x {cs}th column "
                         f"({ref.shape[0] * ref.shape[1]} pixels, {secs:.1f} s)",
            "gpu_rmse": rm(gpu), "fp64_oracle_rmse": rm(f64), "spp": opts.samples}


def frame_leg(name: str, steps: int = 3, warmup: int = 1):
    """Another workload's frames, timed as the headline is (device-resident, wall clock around the frames)."""
    import torch
    from rayna_amd import Renderer

    scene, cam, opts = workload(name)
    with Renderer(scene, cam, opts) as r:
        for _ in range(warmup):
            r.render(to_host=False)
        torch.cuda.synchronize()
        ext = sh = 0.0
        segs = 0
        with _NoGc():
            t0 = time.perf_counter()
            for _ in range(steps):
                st = r.render(to_host=False).stats
                ext += st.extend_ms
                sh += st.shade_ms
                segs += st.segments
            torch.cuda.synchronize()
            secs = time.perf_counter() - t0
    n = opts.width * opts.height * opts.samples
    return {"workload": name, "width": opts.width, "height": opts.height, "spp": opts.samples,
            "ray_depth": opts.ray_depth, "msamples_per_s": n * steps / secs / 1e6, "mrays_per_s": segs / secs / 1e6,
            "ms_per_frame": secs * 1e3 / steps, "extend_ms_per_frame": ext / steps,
            "shade_ms_per_frame": sh / steps, "frames": steps}


def inverse_leg():
    """The inverse-rendering configuration run once at its configured size (diff.recover_colours defaults:
    256x256, a 1024-spp target, 16 spp forward + path-replay adjoint per step, Adam lr 0.02, 200 steps)."""
    from rayna_amd import diff
    res = diff.recover_colours()
    return {"steps": len(res.losses), "loss_first": res.losses[0], "loss_last": res.losses[-1],
            "param_rmse_first": res.param_rmse[0], "param_rmse_last": res.param_rmse[-1],
            "identifiable_colours": res.identifiable, "adjoint_forward_ratio": res.adjoint_forward_ratio,
            "seconds": res.seconds}


def adjoint_leg(steps: int):
    """The inverse-rendering configuration (256x256, 16 spp, the 521 hidden colours of inverse_spheres): device
    time of a forward frame and of its path-replay adjoint, reusing the frame's radiance as the primal pass,
    and of the adjoint when it recomputes the primal pass too.  Medians over `steps` frames."""
    import numpy as np
    from dataclasses import replace
    from rayna_amd import Renderer, presets

    scene, cam, opts, _ = presets.inverse_spheres()
    d = np.random.default_rng(0).standard_normal((opts.height, opts.width, 3)).astype(np.float32)
    fwd, adj, adj_full = [], [], []
    with Renderer(scene, cam, opts) as r:
        for i in range(steps + 1):
            r.set_options(replace(opts, seed=100 + i))
            f = r.render(to_host=False).stats
            _, a = r.adjoint(d, frame=0, layout="rowmajor")
            _, a2 = r.adjoint(d, frame=0, layout="rowmajor")
            if i:  # the first round is warm-up
                fwd.append(f.kernel_ms)
                adj.append(a.kernel_ms)
                adj_full.append(a2.kernel_ms)
    f_ms, a_ms, a2_ms = (float(np.median(x)) for x in (fwd, adj, adj_full))
    return {"workload": "inverse_spheres", "width": opts.width, "height": opts.height, "spp": opts.samples,
            "colours": len(scene.materials_list()), "forward_ms": f_ms, "adjoint_ms": a_ms,
            "adjoint_forward_ratio": a_ms / f_ms, "adjoint_with_primal_ms": a2_ms,
            "adjoint_with_primal_forward_ratio": a2_ms / f_ms}


def batch_leg(tune: bool = True):
    """The batched-rendering configuration: 64 requests (log-uniform 128^2 .. 1920x1080, 1..256 spp, scenes
    drawn from random_spheres / triangle_mesh / sdf_volume, seeds 100..163).  The launch shape comes from the
    autotuner, run on the same requests at 1/2 of the samples, and serves both sides: one render call per
    request (a Renderer per scene, set_camera + set_options + render, in request order) and one batch call
    (with the tuned schedule, by scene or interleaved, and again shortest first; its chunks traced on the tuned
    number of concurrent pipelines).  Throughput from wall time; latency of request i: its completion time
    from the start."""
    import numpy as np
    from rayna_amd import Renderer, presets
    from rayna_amd import autotune as AT
    from rayna_amd.batch import BatchRenderer

    scenes, reqs = presets.batch_requests()
    total = sum(q.opts.width * q.opts.height * q.opts.samples for q in reqs)
    warm = AT.scaled_requests(reqs, 1 << 16)  # one sample each: allocates the largest buffers up front

    half = AT.scaled_requests(reqs, 2)
    cfg = dict(AT.DEFAULTS)
    tuned = None
    if tune:
        with BatchRenderer(scenes) as b:
            b.set_tuning(schedule="by_scene")
            b.render(warm)
            tr = AT.tune(b, half, repeats=1)
        cfg = dict(tr.best)
        tuned = {"best": tr.best, "best_ms": tr.best_ms, "default_ms": tr.default_ms, "trials": len(tr.trials)}
        _progress(f"batch: autotuned ({len(tr.trials)} trials)")
    shape = {k: cfg[k] for k in ("chunk_slots", "refill_lanes", "wave_batch", "shade_blocks_per_cu")}

    firsts = {q.scene: q for q in reversed(reqs)}
    rends = {k: Renderer(scenes[k], q.camera, q.opts) for k, q in firsts.items()}
    for r in rends.values():
        r.set_tuning(**shape)

    def one_by_one(requests):
        lat = []
        with _NoGc():
            t0 = time.perf_counter()
            for q in requests:
                r = rends[q.scene]
                r.set_camera(q.camera)
                r.set_options(q.opts)
                r.render(layout=q.layout)
                lat.append(time.perf_counter() - t0)
            total_s = time.perf_counter() - t0
        return total_s, np.array(lat) * 1e3

    one_by_one(warm)
    seq_runs = []
    for i in range(3):
        seq_runs.append(one_by_one(reqs))
        _progress(f"batch: one request per launch, run {i + 1}: {seq_runs[-1][0]:.2f} s")
    seq_runs.sort(key=lambda x: x[0])
    seq_s, seq_lat = seq_runs[1]  # the median of three
    for r in rends.values():
        r.close()

    # the batch's path state is allocated (by a run at half the samples, whose chunks are as large) only once the
    # renderers have been closed, so that the two sides never hold their buffers at once
    b = BatchRenderer(scenes)
    b.set_tuning(schedule=cfg["schedule"], use_graph=False, pipelines=cfg["pipelines"], **shape)
    b.render(half, to_host=False)

    def batched(schedule):
        b.set_tuning(schedule=schedule, use_graph=False, pipelines=cfg["pipelines"], **shape)  # eager: timed
        with _NoGc():
            t0 = time.perf_counter()
            _, st = b.render(reqs)
            total_s = time.perf_counter() - t0
        return total_s, st

    bat_s, st = batched(cfg["schedule"])
    sjf_s, st2 = batched("shortest_first")
    _progress(f"batch: batched {bat_s:.2f} s, shortest first {sjf_s:.2f} s")
    # the whole launch sequence captured in a hipGraph (the interleaved schedule is traced eager only): the
    # first call captures, the next replays
    graph_schedule = "by_scene" if cfg["schedule"] == "interleaved" else cfg["schedule"]
    b.set_tuning(schedule=graph_schedule, use_graph=True, pipelines=cfg["pipelines"], **shape)
    b.render(reqs, to_host=False)
    with _NoGc():
        t0 = time.perf_counter()
        b.render(reqs)
        graph_s = time.perf_counter() - t0
    b.set_tuning(schedule=graph_schedule, use_graph=False, pipelines=cfg["pipelines"], **shape)
    eager_same_s, _ = batched(graph_schedule)
    b.close()

    def pct(a, q):
        return float(np.percentile(a, q))
    return {"workload": "batch_requests", "requests": len(reqs), "samples": total,
            "one_request_per_launch_msamples_per_s": total / seq_s / 1e6,
            "one_request_per_launch_runs_s": [x[0] for x in seq_runs],
            "graph_replay": {"schedule": graph_schedule, "msamples_per_s": total / graph_s / 1e6, "s": graph_s,
                             "eager_same_schedule_s": eager_same_s},
            "batched_msamples_per_s": total / bat_s / 1e6, "batched_over_one_per_launch": seq_s / bat_s,
            "one_request_per_launch_s": seq_s, "batched_s": bat_s, "batched_kernel_ms": st.kernel_ms,
            "chunks": st.chunks, "launches": st.launches,
            "batched_shortest_first_msamples_per_s": total / sjf_s / 1e6,
            "latency_ms": {"one_request_per_launch": {"p50": pct(seq_lat, 50), "p99": pct(seq_lat, 99)},
                           "batched": {"p50": pct(st.request_ms, 50), "p99": pct(st.request_ms, 99)},
                           "batched_shortest_first": {"p50": pct(st2.request_ms, 50),
                                                      "p99": pct(st2.request_ms, 99)}},
            "schedule": cfg["schedule"], "launch_shape": {**shape, "pipelines": cfg["pipelines"]}, "tuning": tuned}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="triangle_mesh")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU legs' threads (0: the host's share, <= 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward time ratio leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-rendering leg")
    ap.add_argument("--no-tune", action="store_true", help="batched leg with the default launch shape")
    ap.add_argument("--no-workloads", action="store_true", help="skip the other workloads' frames")
    ap.add_argument("--no-inverse", action="store_true", help="skip the inverse-rendering run")
    ap.add_argument("--valu-peak", action="store_true",
                    help="run only the VALU issue-rate calibration kernel (for counter runs) and print its rate")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-rendering leg (for kernel traces) and print its JSON object")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only copy a buffer of this many MiB three times (the counter calibration kernel)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (the renderer is single-GPU; --gpus {a.gpus} is not supported)",
              file=sys.stderr)
        return 2

    import torch
    if a.copy_mib:
        src = torch.ones(a.copy_mib * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copied {a.copy_mib} MiB x3", flush=True)
        return 0
    if a.batch_only:
        print(json.dumps(batch_leg(tune=not a.no_tune)), flush=True)
        return 0
    if a.valu_peak:
        from rayna_amd.renderer import measure_valu_peak
        rate, secs = measure_valu_peak()
        print(f"VALU calibration: {rate / 1e9:.1f} G wave64 instructions/s ({secs * 1e3:.3f} ms)", flush=True)
        return 0
    from rayna_amd import Renderer

    if torch.cuda.is_available():  # the framework's device context first, so that its setup is not timed
        torch.cuda.synchronize()
    scene, cam, opts = workload(a.workload)
    _progress(f"{a.workload}: scene described")
    t_build = time.time()
    r = Renderer(scene, cam, opts)
    t_build = time.time() - t_build
    _progress(f"{a.workload}: renderer created in {t_build:.1f} s")
    for _ in range(a.warmup):
        r.render(to_host=False)
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    ext_ms = sh_ms = kern_ms = call_ms = cam_ms = 0.0
    segs = 0
    launches = 0
    step_ms = []
    nogc = _NoGc()
    nogc.__enter__()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        st = r.render(to_host=False).stats
        step_ms.append((time.perf_counter() - ts) * 1e3)
        ext_ms += st.extend_ms
        cam_ms += st.extend_camera_ms
        sh_ms += st.shade_ms
        kern_ms += st.kernel_ms
        call_ms += st.duration * 1e3
        segs += st.segments
        launches += st.extend_launches
    torch.cuda.synchronize() if torch.cuda.is_available() else None
    t1 = time.perf_counter()
    # the same frames with the image copied to the caller's host buffer (Renderer::render returns the image)
    t2 = time.perf_counter()
    for _ in range(a.steps):
        r.render(layout="xmajor")
    t3 = time.perf_counter()
    nogc.__exit__(None)
    secs = t1 - t0
    _progress(f"{a.workload}: {a.steps} timed frames, {secs * 1e3 / a.steps:.2f} ms each")
    samples_per_frame = opts.width * opts.height * opts.samples
    value = samples_per_frame * a.steps / secs / 1e6
    ext_bytes = EXTEND_BYTES_PER_SEGMENT * segs
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    per_seg = EXTEND_FABRIC_BYTES_PER_SEGMENT.get(a.workload)
    traffic = per_seg * segs / launches if per_seg and launches else None
    frame_bytes = FRAME_BYTES_PER_SEGMENT * segs + FRAME_BYTES_PER_SAMPLE * samples_per_frame * a.steps
    line = {
        "metric": "Msamples/s at 1080p (triangle-mesh workload, equal-seed parity with the fp32 CPU restatement)",
        "value": value,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": secs * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": a.workload, "width": opts.width, "height": opts.height, "spp": opts.samples,
                   "ray_depth": opts.ray_depth, "triangles": scene.n_triangles(), "seed": opts.seed},
        "mrays_per_s": segs / secs / 1e6,
        "segments_per_frame": segs // max(1, a.steps),
        "extend_ms_per_frame": ext_ms / a.steps,
        "extend_camera_ms_per_frame": cam_ms / a.steps,
        "shade_ms_per_frame": sh_ms / a.steps,
        "kernel_ms_per_frame": kern_ms / a.steps,
        "render_call_ms_per_frame": call_ms / a.steps,
        "step_ms_min_median_max": [min(step_ms), sorted(step_ms)[len(step_ms) // 2], max(step_ms)],
        "extend_launches_per_frame": launches // max(1, a.steps),
        "frame_alg_gbps": frame_bytes / secs / 1e9,
        "to_host": {"msamples_per_s": samples_per_frame * a.steps / (t3 - t2) / 1e6,
                    "ms_per_step": (t3 - t2) * 1e3 / a.steps,
                    "note": "each frame's image copied to host memory (x-major float32), as Renderer::render hands it"},
        "scene_build_s": t_build,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_kind": "L2-miss (fabric) bytes per launch: rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "calibrated on a 256 MiB copy; Infinity Cache hits included",
                     "kernel": "extend: k_extend_packet (camera rays) + k_extend (later depths)",
                     "bytes_per_segment": EXTEND_BYTES_PER_SEGMENT},
    }
    valu = valu_roofline(a.workload, samples_per_frame * a.steps, segs, ext_ms, cam_ms)
    if valu:
        valu["k_extend_ms_per_frame"] = (ext_ms - cam_ms) / a.steps
        line["roofline"]["valu"] = valu
    threads, host = host_cpus()
    if a.cpu_threads:
        threads = a.cpu_threads
    tw = None
    if not a.no_cpu_baseline:
        import oracle
        tw = oracle.TwinScene(scene)
        _progress("oracle scene built")
        r.clear_accumulation()
        line["rmse"] = rmse_leg(r, tw, cam, opts, threads)
        _progress("rmse leg done")
    r.close()
    if not a.no_workloads:
        line["workloads"] = {}
        for w in ("random_spheres", "sdf_volume", "triangle_mesh", "rttnw", "cornell"):
            if w != a.workload:
                line["workloads"][w] = frame_leg(w)
                _progress(f"workload {w} done")
    if not a.no_adjoint:
        line["adjoint"] = adjoint_leg(max(3, a.steps))
        _progress("adjoint leg done")
    if not a.no_inverse:
        line["inverse_rendering"] = inverse_leg()
        _progress("inverse rendering done")
    if not a.no_batch:
        line["batch"] = batch_leg(tune=not a.no_tune)
        _progress("batch leg done")
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(tw, cam, opts, a.cpu_seconds, threads, host)
        _progress("cpu baseline done")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
