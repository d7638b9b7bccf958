"""Benchmark of the plain delta-tracking volumetric path tracer on one MI355X.

Workload (sparse512): sparse 512^3 NanoVDB-layout fBm cloud (25 % of 8^3 leaves active, sigma_t max 50),
albedo 0.999, Henyey-Greenstein g = 0.85, Hosek-Wilkie sky (turbidity 3, sun elevation 30 deg) sampled
with alias tables, 1920x1080, up to 256 bounces, Russian roulette after bounce 8.  One step = one frame
of the workload: 16 samples per pixel over the full image in one launch.

Prints one JSON line: value = path vertices per second of the plain tracer (rays traced through the
medium per second, all inputs resident in HBM) with local-majorant tracking (a walk over the grid's
32^3-voxel regions, their majorants staged in LDS), with the trace kernel's roofline (HIP-event kernel
time, 32 algorithmic bytes per trilinear density lookup) and the CPU oracle timed on the host's cores as
a baseline.  The reference's own global-majorant tracking of the same frame is a leg beside it.  Two NRC legs ride along: the standalone network step (2^16 training samples + 2^21 inference
queries, 5 hidden layers, FP16) giving nrc_queries_per_s, and NRC-integrated rendering of the same scene.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "path vertices/sec + NRC queries/sec; wall time to equal RMSE vs FP64 oracle"
HBM_PEAK_GBS = 8000.0
BYTES_PER_LOOKUP = 32  # 8 FP32 voxel taps per trilinear lookup
FP16_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
FP8_PEAK_TFLOPS = 5000.0  # MI355X dense FP8 MFMA
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 MFMA
# NRC settings of the equal-RMSE leg: the reference's network (hash grid, 2 hidden layers, lr 1e-2) with
# its radiance-scale knob (renderer_main.cpp:979-995) set so the cloud's radiance (~0.03-0.3) sits well
# above the loss's 0.01 epsilon, and a shorter weight EMA than the reference's 0.99 (the benchmark
# frame has only ~44 optimiser steps, over which a 0.99 average still lags the training weights)
NRC_RADIANCE_SCALE = 10.0
NRC_EMA_DECAY = 0.95
# initial loss scales of the equal-RMSE legs: tiny-cuda-nn's 128, with which every training step is
# applied, and 2^22, with which the fresh network's first steps (their loss gradients hold outliers that
# overflow FP16 at this scale) are skipped until the dynamic scale has backed off
NRC_LOSS_SCALES = (("", 128.0), ("_scale2^22", 2.0 ** 22))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sparse512")
    ap.add_argument("--tracking", choices=["local", "global"], default="local")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nrc-steps", type=int, default=8, help="steps of each NRC leg (0 skips them)")
    ap.add_argument("--rmse-ref-spp", type=int, default=256,
                    help="spp of each half of the equal-RMSE reference (0 skips the equal-RMSE leg)")
    ap.add_argument("--crop-ref-spp", type=int, default=256,
                    help="spp of each half of the FP64 oracle's 64 x 64 crop judging the equal-RMSE leg (0 skips it)")
    ap.add_argument("--legs", choices=["all", "none", "network", "dense128"], default="all",
                    help="legs beside the headline frame: all of them, none (e.g. for counter passes), only the NRC "
                         "network steps, or only the equal-RMSE leg on the dense 128^3 cloud (e.g. to profile the NRC "
                         "schedule)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (HBM counter calibration)")
    return ap.parse_args()


def workload_text(cfg, spp: int) -> str:
    density = ("density (g - thr) / (g_max - thr), contrast 1" if cfg.fbm.contrast == 1.0 else
               f"density min({cfg.fbm.contrast:g} (g - thr) / (g_max - thr), 1)")
    return (f"{cfg.name}: sparse {cfg.fbm.resolution}^3 thresholded fBm cloud ({density}), 25% of 8^3 leaves, "
            f"sigma_t max {cfg.density_coeff:g}, albedo {cfg.albedo}, HG g={cfg.forwardness}, Hosek-Wilkie T=3 sun "
            f"30deg, alias-table env sampling, {cfg.width}x{cfg.height}, {spp} spp per step, <= "
            f"{cfg.max_path_length} bounces, RR after {cfg.rr_after_path_length}")


def saturated_leg(args, stream, tracking: int) -> dict:
    """The plain tracer on the same cloud with contrast 4 (a saturated interior at the majorant), with the
    given tracking mode."""
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.device import EventTimer
    from rtcamp8_amd.tracer import PathTracer

    cfg = sc.CONFIGS["sparse512_saturated"].with_(tracking=tracking)
    scene = sc.DeviceScene(cfg, stream.handle)
    pt = PathTracer(scene, stream=stream.handle)
    params = pt.params()
    pt.trace(cfg.spp, params)
    stream.synchronize()
    pt.stats.fill(0)
    steps = max(2, args.steps // 2)
    timers = [EventTimer() for _ in range(steps)]
    t0 = time.perf_counter()
    for t in timers:
        t.start(stream.handle)
        pt.trace(cfg.spp, params)
        t.stop(stream.handle)
    stream.synchronize()
    wall = time.perf_counter() - t0
    c = pt.counters()
    ms = float(np.mean([t.elapsed_ms() for t in timers]))
    for t in timers:
        t.destroy()
    out = {"workload": workload_text(cfg, cfg.spp), "steps": steps, "ms_per_step": wall * 1e3 / steps,
           "kernel_ms_avg": ms, "path_vertices_per_s": c["vertices"] / wall, "paths_per_s": c["paths"] / wall,
           "vertices_per_path": c["vertices"] / max(c["paths"], 1), "lookups_per_path": c["lookups"] / max(c["paths"], 1),
           "roofline_frac": c["lookups"] / steps * BYTES_PER_LOOKUP / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    pt.close()
    scene.close()
    out["tracking"] = "local" if tracking else "global"
    print(f"[bench] saturated cloud ({out['tracking']} majorant): {out['path_vertices_per_s']:.3g} vertices/s, "
          f"{ms:.2f} ms per frame", file=sys.stderr, flush=True)
    return out


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary (profiles/<kernel>_<workload>_hbm.json),
    if there is one."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_hbm.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(scene, target_s: float):
    """The CPU oracle (FP32 restatement of the same algorithm) on all host cores, over full-frame passes
    at 1 spp until about `target_s` seconds of work."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import download_env, download_grid, oracle, oracle_seed, ptr  # oracle: checker / baseline only

    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    cfg = scene.cfg
    hg = download_grid(scene.grid)
    he, keep = download_env(scene.env)
    p = sc.fill_params(cfg, hg.grid, he, scene.camera)
    w, h = cfg.width, cfg.height
    rng = oracle_seed(w, h, cfg.rng_seed)
    accum = np.zeros(3 * w * h, dtype=np.float64)
    stats = np.zeros(_abi.STAT_COUNT, dtype=np.uint64)
    p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), ptr(accum), None, ptr(stats)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # full-frame passes at 1 spp each (the RNG states advance between passes), the last one cut to a
    # centred band of rows so the sample lands near target_s
    dt, passes, rows_done = 0.0, 0, 0
    rows = h
    while dt < target_s and passes < 64:
        y0 = (h - rows) // 2
        t0 = time.perf_counter()
        assert oracle().oracle_trace_f32(C.byref(p), _abi.ENTRY_PATH_TRACE, 1, y0, y0 + rows, threads) == 0
        dt += time.perf_counter() - t0
        passes += 1
        rows_done += rows
        rate = rows_done / max(dt, 1e-3)
        rows = int(min(h, max(16, rate * (target_s - dt))))
    vertices = int(stats[_abi.STAT_VERTICES])
    return {"value": vertices / dt, "unit": "path vertices/s", "cores": threads, "kind": "port",
            "sample": f"{passes} passes over {rows_done} image rows of {w} px ({rows_done * w} paths at 1 spp each), "
                      f"{vertices} vertices in {dt:.1f} s, FP32 oracle",
            "seconds": dt}


def random_leaf_queries(scene, n: int, seed: int) -> np.ndarray:
    """NRC queries uniform over the volume's active leaves: position = normalised index-space point inside
    a random active leaf, directions uniform, albedo as the medium's."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.device import download_ptr

    g = scene.grid
    origins = download_ptr(g.leaf_origins, np.int32, 3 * g.num_leaves).reshape(-1, 3)
    lo = np.array(g.index_bbox_min[:], dtype=np.float64)
    ext = np.array(g.index_bbox_max[:], dtype=np.float64) + 1.0 - lo
    rng = np.random.default_rng(seed)
    leaf = origins[rng.integers(0, origins.shape[0], n)]
    pos = (leaf + rng.random((n, 3)) * 8.0 - lo) / ext
    q = np.zeros((n, 14), dtype=np.float32)
    q[:, 0:3] = np.clip(pos, 0.0, 1.0)
    q[:, 3] = np.arccos(1.0 - 2.0 * rng.random(n))
    q[:, 4] = rng.random(n) * 2.0 * np.pi
    q[:, 5:7] = q[:, 3:5]
    q[:, 8:11] = scene.cfg.albedo
    return q


def nrc_mlp_leg(scene, steps: int, hidden_layers: int = 5, n_infer: int = 1 << 21, n_train: int = 1 << 16,
                encoding: int = 0, precision: int | None = None):
    """Standalone network step: train on 2^16 samples, then infer 2^21 queries (FP16 storage, FP32
    accumulation; BF16 training with E4M3 inference; or FP32 throughout); the benchmark configuration is 5 hidden layers
    of 64 with the frequency (triangle wave) encoding, the renderer's default 2 hidden layers with the
    hash grid."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.device import DeviceBuffer, EventTimer
    from rtcamp8_amd.nrc import NeuralRadianceCache

    q = random_leaf_queries(scene, n_infer, seed=1234)
    t = (0.2 + np.random.default_rng(99).random((n_train, 3))).astype(np.float32)
    dq, dt, dout = DeviceBuffer.from_array(q), DeviceBuffer.from_array(t), DeviceBuffer(12 * n_infer)
    cache = NeuralRadianceCache()
    precision = _abi.NRC_FP16 if precision is None else precision
    label, peak = {_abi.NRC_FP16: ("fp16 storage, fp32 accumulate", FP16_PEAK_TFLOPS),
                   _abi.NRC_BF16: ("bf16 storage, fp32 accumulate", FP16_PEAK_TFLOPS),
                   _abi.NRC_BF16_FP8: ("bf16 training, e4m3 inference, fp32 accumulate", FP8_PEAK_TFLOPS),
                   _abi.NRC_FP32: ("fp32 throughout", FP32_PEAK_TFLOPS)}[precision]
    cache.initialize(num_hidden_layers=hidden_layers, learning_rate=1e-2, precision=precision, encoding=encoding)
    for _ in range(2):
        cache.train(None, dq.ptr, dt.ptr, n_train)
        cache.infer(None, dq.ptr, n_infer, dout.ptr)
    tt, ti = [EventTimer() for _ in range(steps)], [EventTimer() for _ in range(steps)]
    for i in range(steps):
        tt[i].start(None)
        cache.train(None, dq.ptr, dt.ptr, n_train)
        tt[i].stop(None)
        ti[i].start(None)
        cache.infer(None, dq.ptr, n_infer, dout.ptr)
        ti[i].stop(None)
    train_ms = float(np.mean([x.elapsed_ms() for x in tt]))
    infer_ms = float(np.mean([x.elapsed_ms() for x in ti]))
    out = dout.download(np.float32, 3 * n_infer)
    ok = bool(np.all(np.isfinite(out)))
    for x in tt + ti:
        x.destroy()
    cache.finalize()
    for b in (dq, dt, dout):
        b.free()
    fwd_flop = 2.0 * (hidden_layers * 64 * 64 + 64 * 3)
    train_flop = 3.0 * fwd_flop  # forward + backward (input and weight gradients)
    return {
        "hidden_layers": hidden_layers,
        "precision": label,
        "encoding": "hash grid (16 levels x 2, 2^15)" if encoding == 1 else "triangle waves (12 frequencies)",
        "infer_queries": n_infer, "train_samples": n_train,
        "infer_ms": infer_ms, "train_ms": train_ms,
        "queries_per_s": n_infer / (infer_ms * 1e-3),
        "train_samples_per_s": n_train / (train_ms * 1e-3),
        "infer_tflops": n_infer * fwd_flop / (infer_ms * 1e-3) / 1e12,
        "train_tflops": n_train * train_flop / (train_ms * 1e-3) / 1e12,
        "mfma_peak_tflops": peak,
        "infer_frac": n_infer * fwd_flop / (infer_ms * 1e-3) / 1e12 / peak,
        "finite": ok,
    }


MULTI_INSTANCES = [  # (fBm resolution, sparse, sigma_t max, image width, height)
    (64, False, 5.0, 256, 256), (96, False, 10.0, 320, 240), (128, False, 20.0, 512, 512),
    (192, True, 40.0, 640, 360), (256, True, 60.0, 800, 600), (384, True, 100.0, 1024, 576),
    (512, True, 150.0, 1280, 720), (1024, True, 200.0, 1920, 1080)]


def multi_instance_leg(steps: int, stream):
    """Uneven multi-instance batch: eight fBm volumes from 64^3 to 1024^3 (sigma_t max 5 to 200) at image
    sizes from 256^2 to 1920x1080, each with its own NRC (hash grid, BF16 training, E4M3 inference),
    after the first frame's training.  One NRC render launch per instance against all eight in one
    persistent launch, with the device's workgroups shared by the autotuner (each instance's own launch
    time) or by pixels."""
    from rtcamp8_amd import _abi, batch
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import EventTimer
    from rtcamp8_amd.nrc import NrcRenderer

    base = sc.CONFIGS["sparse512"]
    scenes, rs = [], []
    for res, sparse, sigma, w, h in MULTI_INSTANCES:
        fbm = sc.FbmConfig(resolution=res, seed=0x5EED0100 + res, sparse=sparse, active_leaf_fraction=0.25,
                           contrast=4.0 if sparse else 1.0)
        cfg = base.with_(name=f"fbm{res}", width=w, height=h, density_coeff=sigma, fbm=fbm)
        scenes.append(sc.DeviceScene(cfg, stream.handle))
        rs.append(NrcRenderer(scenes[-1], num_hidden_layers=2, precision=_abi.NRC_BF16_FP8, stream=stream.handle,
                              radiance_scale=NRC_RADIANCE_SCALE, ema_decay=NRC_EMA_DECAY, samples_per_launch=1))
    for r in rs:
        r.render_frame(spp=1)
    stream.synchronize()

    def params(r):
        p = r.params()
        p.nrc.num_samples = 1
        return p

    def one(i):
        call("rtc8x_trace", C.c_void_p(stream.handle), C.byref(params(rs[i])), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1)

    def one_with(i, ev, shape):
        call("rtc8x_trace", C.c_void_p(stream.handle), C.byref(batch.apply_knobs(params(rs[i]), (ev, shape))),
             _abi.ENTRY_PATH_TRACE_WITH_NRC, 1)

    def timed(fn):
        fn()
        ms = []
        for _ in range(steps):
            t = EventTimer()
            t.start(stream.handle)
            fn()
            t.stop(stream.handle)
            stream.synchronize()
            ms.append(t.elapsed_ms())
            t.destroy()
        return float(np.median(ms))

    ones = [1] * len(rs)
    seq_ms = timed(lambda: [one(i) for i in range(len(rs))])
    pixel_ms = timed(lambda: batch._launch([params(r) for r in rs], ones, _abi.ENTRY_PATH_TRACE_WITH_NRC, None,
                                           stream.handle))
    t_tune = time.perf_counter()
    solo_shares, solo_knobs = batch.autotune(one_with, len(rs), stream=stream.handle)
    default_knobs = [(0, 0)] * len(rs)

    def launch_batch(sh, kn):
        for r, k in zip(rs, kn):
            r.knobs = k
        batch._launch([params(r) for r in rs], ones, _abi.ENTRY_PATH_TRACE_WITH_NRC, sh, stream.handle)

    settings = [(solo_shares, solo_knobs), (None, solo_knobs), (solo_shares, default_knobs), (None, default_knobs)]
    shares, knobs, _ = batch.pick_batch_setting(launch_batch, settings, stream=stream.handle)
    tune_s = time.perf_counter() - t_tune
    for r, k in zip(rs, solo_knobs):
        r.knobs = k
    seq_tuned_ms = timed(lambda: [one(i) for i in range(len(rs))])
    for r, k in zip(rs, knobs):
        r.knobs = k
    tuned_ms = timed(lambda: batch._launch([params(r) for r in rs], ones, _abi.ENTRY_PATH_TRACE_WITH_NRC, shares,
                                           stream.handle))
    choice = {"shares": "pixels" if shares is None else "solo launch times",
              "knobs": "defaults" if knobs is default_knobs else "tuned per instance"}
    batch.render_nrc(rs, ones, shares, stream.handle)
    stream.synchronize()
    finite = all(bool(np.all(np.isfinite(r.image()))) for r in rs)
    for r in rs:
        r.close()
    for s_ in scenes:
        s_.close()
    return {"instances": [{"fbm_resolution": res, "sparse": sparse, "sigma_t_max": sigma, "image": [w, h],
                           "own_launch_ms": t}
                          for (res, sparse, sigma, w, h), t in zip(MULTI_INSTANCES, solo_shares)],
            "nrc": "hash grid, 2 hidden layers, bf16 training, e4m3 inference; first frame (8 training iterations) done",
            "entry": "PATH_TRACE_WITH_NRC, 1 spp per instance",
            "one_launch_per_instance_ms": seq_ms, "one_launch_per_instance_tuned_ms": seq_tuned_ms,
            "batched_autotuned_ms": tuned_ms, "batched_by_pixels_ms": pixel_ms,
            "autotuned": [{"event_eighths": k[0], "tile": batch.TILE_SHAPES[k[1]], "share_ms": sh}
                          for k, sh in zip(solo_knobs, solo_shares)],
            "batch_choice": choice,
            "autotune_seconds": tune_s,
            "speedup_autotuned": seq_ms / tuned_ms, "finite": finite}


def nrc_render_leg(scene, steps: int, stream):
    """NRC-integrated rendering of the same scene: per step one training iteration and one sample per
    pixel traced into the cache (2 hidden layers, FP16), after the first frame's 8 training iterations."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.device import EventTimer
    from rtcamp8_amd.nrc import NrcRenderer

    r = NrcRenderer(scene, num_hidden_layers=2, stream=stream.handle, train_tracking=_abi.TRACKING_GLOBAL_MAJORANT)
    r.render_frame(spp=1)
    stream.synchronize()
    r.stats_buf.fill(0)
    q0 = r.stats.inference_queries
    ev = EventTimer()
    ev.start(stream.handle)
    t0 = time.perf_counter()
    for _ in range(steps):
        r.render_frame(spp=1, train_iterations=1)
    stream.synchronize()
    wall = time.perf_counter() - t0
    ev.stop(stream.handle)
    gpu_ms = ev.elapsed_ms()
    ev.destroy()
    c = r.counters()
    img = r.image()
    res = {
        "ms_per_spp": wall * 1e3 / steps, "gpu_ms_per_spp": gpu_ms / steps,
        "spp": steps, "hidden_layers": 2,
        "inference_queries_per_s": (r.stats.inference_queries - q0) / wall,
        "path_vertices_per_s": c["vertices"] / wall,
        "vertices_per_path": c["vertices"] / max(c["paths"], 1),
        "finite": bool(np.all(np.isfinite(img))),
        "last_loss": r.stats.last_loss,
    }
    r.close()
    return res


def latency_roofline(scene, params, lookups_per_launch: float, lookups_per_path: float, trace_lookups_per_s: float,
                     paths_per_s: float, waves=(3, 4)):
    """Memory-latency roofline of the tracer's density lookups: the march probe (rtc8x_march_probe) walks
    the tracer's own stepping pattern on this frame -- persistent lanes taking the frame's pixels in 8x8
    tiles, the same camera rays and PCG states, delta tracking at the majorant's spacing with two tentative
    collisions looked up together, a ratio-tracked shadow segment towards the sun and a phase sample at
    every real collision -- with none of the tracer's vertex shading, and every lane stops at the same
    lookup budget (as many lookups in all as one traced frame) so the launch has no tail.  The reference
    point is the tracer's occupancy (3 waves per SIMD, registers allotted as for the tracer), 4 waves
    beside it; the gather probe (random positions of random leaves, no coherence) is reported too."""
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer, EventTimer, device_cus

    g = scene.grid
    light = (C.c_float * 3)(*[float(v) for v in sc.sun_direction(scene.cfg)])
    stats = DeviceBuffer.zeros(16)
    cus = device_cus()
    walk = ("the local tracker's region walk over the LDS-staged region table, one tentative collision per lookup round"
            if params.tracking else "delta tracking at the global majorant's spacing, 2 tentative collisions per lookup round")
    res = {"probe": "march probe: persistent lanes walk the frame's camera paths in 8x8 pixel tiles, "
                    f"{scene.cfg.spp} per pixel ({walk}, shadow segment towards the sun and a phase sample per "
                    "collision, no vertex shading), a fixed lookup budget per lane totalling one frame's lookups",
           "tracking": "local" if params.tracking else "global"}
    for w in waves:
        budget = max(1, int(lookups_per_launch / (cus * w * 256)))
        call("rtc8x_march_probe", None, C.byref(params), light, scene.cfg.spp, budget, w, C.c_void_p(stats.ptr))
        stats.fill(0)
        t = EventTimer()
        t.start(None)
        call("rtc8x_march_probe", None, C.byref(params), light, scene.cfg.spp, budget, w, C.c_void_p(stats.ptr))
        t.stop(None)
        ms = t.elapsed_ms()
        t.destroy()
        looked, paths = (int(v) for v in stats.download(np.uint64, 2))
        res[f"lookups_per_s_{w}waves"] = looked / (ms * 1e-3)
        res[f"probe_ms_{w}waves"] = ms
        res[f"budget_per_lane_{w}waves"] = budget
        res[f"lookups_per_path_{w}waves"] = looked / max(paths, 1)
    stats.free()
    ref = res[f"lookups_per_s_{waves[0]}waves"]
    res.update({"waves_per_simd": waves[0], "trace_lookups_per_s": trace_lookups_per_s,
                "frac": trace_lookups_per_s / ref,
                "paths_per_s": paths_per_s, "roofline_paths_per_s": ref / lookups_per_path,
                "paths_frac": paths_per_s / (ref / lookups_per_path)})
    # the random-gather probe: the tracer's lookup at random positions of random leaves
    threads, steps = 1920 * 1088, 32
    out = DeviceBuffer(4 * threads)
    for width in (1, 2):
        call("rtc8x_gather_probe", None, C.byref(g), threads, steps, width, waves[0], 7, C.c_void_p(out.ptr))
        t = EventTimer()
        t.start(None)
        call("rtc8x_gather_probe", None, C.byref(g), threads, steps, width, waves[0], 8, C.c_void_p(out.ptr))
        t.stop(None)
        res[f"random_gather_lookups_per_s_x{width}"] = threads * steps * width / (t.elapsed_ms() * 1e-3)
        t.destroy()
    out.free()
    return res


def fp64_crop(scene, spp: int, seed: int, crop) -> np.ndarray:
    """The FP64 CPU oracle's plain-tracer image of the pixels crop = (x0, x1, y0, y1) of the frame."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import download_env, download_grid, oracle, oracle_seed, ptr  # oracle: the judge only

    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    cfg = scene.cfg.with_(tracking=_abi.TRACKING_GLOBAL_MAJORANT)
    hg = download_grid(scene.grid)
    he, keep = download_env(scene.env)
    p = sc.fill_params(cfg, hg.grid, he, scene.camera)
    rng = oracle_seed(cfg.width, cfg.height, seed)
    accum = np.zeros(3 * cfg.width * cfg.height)
    p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), ptr(accum), None, None
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    x0, x1, y0, y1 = crop
    assert oracle().oracle_trace_window_f64(C.byref(p), spp, x0, x1, y0, y1, threads) == 0
    return accum.reshape(cfg.height, cfg.width, 3)[y0:y1, x0:x1] / spp


def equal_rmse_leg(scene, ref_spp: int, spp: int, stream, precisions=None, encoding=None, crop_ref_spp: int = 0):
    """Wall time to equal RMSE: NRC-accelerated rendering vs the plain tracer on the same frame.

    NRC runs the reference's batch schedule (training iterations per frame, then the frame's samples) with
    4 spp per frame.  The full-frame judge is the plain tracer at 2 x ref_spp (two independent halves A, B);
    the MSE of an estimate X is mean((X - A)(X - B)), free of the judge's noise.  The plain tracer's MSE falls
    as 1/spp, so it needs n* = spp * mse_plain / mse_nrc samples to match the NRC image; it is then run at n*
    (rounded up) and its measured time and error are reported.  With crop_ref_spp > 0, a 64 x 64 crop at the
    image centre is also judged against the FP64 CPU oracle at 2 x crop_ref_spp."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.convergence import cross_mse, nrc_image, plain_image

    if precisions is None:
        precisions = (("fp16", _abi.NRC_FP16), ("bf16", _abi.NRC_BF16), ("bf16_fp8", _abi.NRC_BF16_FP8))
    runs = [(name + suffix, prec, ls) for suffix, ls in NRC_LOSS_SCALES for name, prec in precisions]
    encoding = _abi.NRC_HASH_GRID if encoding is None else encoding
    cfg = scene.cfg
    a, ta = plain_image(scene, ref_spp, seed=0xA11CE, stream=stream.handle)
    b, _ = plain_image(scene, ref_spp, seed=0xB0B, stream=stream.handle)
    x, tx = plain_image(scene, spp, seed=0xC0FFEE, stream=stream.handle)
    mean = float((a + b).mean() / 2)
    mse_plain = cross_mse(x, a, b)
    crop = None
    if crop_ref_spp > 0:
        cx, cy = cfg.width // 2, cfg.height // 2
        crop = (cx - 32, cx + 32, cy - 32, cy + 32)
        t0 = time.perf_counter()
        # seeds of their own: the FP64 halves must be independent of every GPU image they judge
        ca, cb = fp64_crop(scene, crop_ref_spp, 0xF64A, crop), fp64_crop(scene, crop_ref_spp, 0xF64B, crop)
        crop_secs = time.perf_counter() - t0
        cmean = float((ca + cb).mean() / 2)

        def sub_rel(sub):
            return float(np.sqrt(max(cross_mse(sub, ca, cb), 0.0)) / cmean)

        def crop_rel(img):
            return sub_rel(img[crop[2]:crop[3], crop[0]:crop[1]])
    res = {"judge": f"plain GPU tracer, 2 x {ref_spp} spp with independent seeds (cross-product MSE)",
           "nrc_schedule": "frame 0: 8 training iterations, later frames 1; 4 spp traced into the cache per frame; "
                           f"the render launches track with the {'local' if cfg.tracking else 'global'} majorant, "
                           "the training launches with the global one",
           "nrc_settings": {"encoding": "hash grid" if encoding == _abi.NRC_HASH_GRID else "triangle waves",
                            "hidden_layers": 2, "radiance_scale": NRC_RADIANCE_SCALE,
                            "ema_decay": NRC_EMA_DECAY},
           "spp": spp, "plain_seconds": tx, "plain_rel_rmse": float(np.sqrt(max(mse_plain, 0.0)) / mean)}
    if crop:
        res["fp64_crop"] = {"crop": list(crop), "oracle_spp": f"2 x {crop_ref_spp}", "oracle_seconds": crop_secs,
                            "plain_rel_rmse": crop_rel(x),
                            # the GPU judge against the FP64 oracle: the judge's own noise, no bias
                            "gpu_judge_rel_rmse": crop_rel((a + b) / 2)}
    for name, prec, ls in runs:
        run = nrc_image(scene, spp, seed=0xD00D, stream=stream.handle, precision=prec, spp_per_frame=4,
                        encoding=encoding, radiance_scale=NRC_RADIANCE_SCALE, ema_decay=NRC_EMA_DECAY,
                        train_tracking=_abi.TRACKING_GLOBAL_MAJORANT, loss_scale=ls)
        mse = cross_mse(run.image, a, b)
        n_star = spp * mse_plain / mse if mse > 0 else float("inf")
        entry = {"initial_loss_scale": ls, "nrc_seconds": run.seconds,
                 "nrc_rel_rmse": float(np.sqrt(max(mse, 0.0)) / mean), "plain_spp_for_equal_rmse": n_star}
        if np.isfinite(n_star) and n_star <= 4096:
            n_run = int(np.ceil(n_star))
            t0 = time.perf_counter()
            y, _ = plain_image(scene, n_run, seed=0x5EED, stream=stream.handle)
            entry.update({"plain_spp_run": n_run, "plain_seconds_measured": time.perf_counter() - t0,
                          "plain_rel_rmse_measured": float(np.sqrt(max(cross_mse(y, a, b), 0.0)) / mean)})
            entry["speedup"] = entry["plain_seconds_measured"] / run.seconds
            if crop:
                entry["fp64_crop"] = {"nrc_rel_rmse": crop_rel(run.image), "plain_rel_rmse_at_n": crop_rel(y)}
        else:
            entry["speedup"] = None
        res[name] = entry
    return res


def dense128_equal_rmse_leg(ref_spp: int, stream, spp: int = 16, crop_ref_spp: int = 0) -> dict:
    """The NRC on the dense 128^3 cloud (256 x 256, <= 16 bounces, constant sky and sun) against the
    plain tracer at equal RMSE, with the network in BF16 and in FP32.  The reference images are the
    plain tracer at 2 x 4 ref_spp; the plain tracer converges to the FP64 oracle on this scene (GPU
    test test_converges_to_fp64_oracle), so they stand in for it."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    small = sc.DeviceScene(sc.CONFIGS["dense128"])
    out = {"workload": "dense128: dense 128^3 fBm, sigma_t max 20, albedo 0.99, HG g=0.8, constant white sky + sun, "
                       f"256x256, <= 16 bounces, {spp} spp"}
    try:
        for enc_name, enc in (("hash_grid", _abi.NRC_HASH_GRID), ("triangle_waves", _abi.NRC_TRIANGLE_WAVE)):
            out[enc_name] = equal_rmse_leg(small, 4 * ref_spp, spp, stream, encoding=enc,
                                           precisions=(("bf16", _abi.NRC_BF16), ("fp32", _abi.NRC_FP32)),
                                           crop_ref_spp=2 * crop_ref_spp)
            r = out[enc_name]
            print(f"[bench] dense128 equal RMSE ({enc_name}): plain rel RMSE {r['plain_rel_rmse']:.4f} in "
                  f"{r['plain_seconds'] * 1e3:.2f} ms; " + "; ".join(
                      f"NRC {k} {r[k]['nrc_rel_rmse']:.4f} in {r[k]['nrc_seconds'] * 1e3:.2f} ms -> "
                      f"{r[k]['speedup'] if r[k]['speedup'] is None else round(r[k]['speedup'], 2)}x"
                      for k in r if isinstance(r[k], dict) and "nrc_rel_rmse" in r[k]), file=sys.stderr, flush=True)
    finally:
        small.close()
    return out


def global_majorant_leg(scene, pt, spp: int, steps: int, stream) -> dict:
    """The same frame traced with the reference's global-majorant tracking (optix_pathtracing_kernels.cu:98-126,
    renderer_kernel_common.h:360-383), bit-exact with the oracle's restatement of it, and its own latency
    roofline."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.device import EventTimer

    params = pt.params()
    params.tracking = _abi.TRACKING_GLOBAL_MAJORANT
    pt.trace(spp, params)
    stream.synchronize()
    pt.stats.fill(0)
    timers = [EventTimer() for _ in range(steps)]
    t0 = time.perf_counter()
    for t in timers:
        t.start(stream.handle)
        pt.trace(spp, params)
        t.stop(stream.handle)
    stream.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [t.elapsed_ms() for t in timers]
    for t in timers:
        t.destroy()
    c = pt.counters()
    lookups_per_launch = c["lookups"] / steps
    ms = float(np.mean(kernel_ms))
    out = {"steps": steps, "ms_per_step": wall * 1e3 / steps, "kernel_ms_avg": ms,
           "path_vertices_per_s": c["vertices"] / wall, "paths_per_s": c["paths"] / wall,
           "vertices_per_path": c["vertices"] / max(c["paths"], 1), "lookups_per_path": c["lookups"] / max(c["paths"], 1),
           "hbm_roofline_frac": lookups_per_launch * BYTES_PER_LOOKUP / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel": "wavefront_kernel<PATH_TRACE, GLOBAL_MAJORANT>"}
    out["latency_roofline"] = latency_roofline(scene, params, lookups_per_launch, out["lookups_per_path"],
                                               c["lookups"] / sum(kernel_ms) * 1e3, out["paths_per_s"])
    pt.stats.fill(0)
    print(f"[bench] global-majorant tracking: {ms:.2f} ms per frame, {out['path_vertices_per_s']:.3g} vertices/s, "
          f"{out['lookups_per_path']:.1f} lookups/path, {out['latency_roofline']['paths_frac']:.2f} of its latency "
          "roofline", file=sys.stderr, flush=True)
    return out


def calibrate_copy(mib: int, steps: int) -> int:
    """HBM counter calibration: `steps` launches of the plain copy kernel over `mib` MiB."""
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer, EventTimer

    nbytes = mib << 20
    a, b = DeviceBuffer.zeros(nbytes), DeviceBuffer(nbytes)
    call("rtc8x_copy_async", C.c_void_p(b.ptr), C.c_void_p(a.ptr), nbytes, None)
    timers = [EventTimer() for _ in range(steps)]
    for t in timers:
        t.start(None)
        call("rtc8x_copy_async", C.c_void_p(b.ptr), C.c_void_p(a.ptr), nbytes, None)
        t.stop(None)
    ms = float(np.mean([t.elapsed_ms() for t in timers]))
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": steps + 1, "ms_avg": ms,
                      "read_plus_write_gbs": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); run it with --gpus 1", file=sys.stderr)
        return 2
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy, args.steps)
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.device import EventTimer, Stream, device_name, synchronize
    from rtcamp8_amd.tracer import PathTracer

    tracking = _abi.TRACKING_LOCAL_MAJORANT if args.tracking == "local" else _abi.TRACKING_GLOBAL_MAJORANT
    cfg = sc.CONFIGS[args.config].with_(tracking=tracking)
    print(f"[bench] device: {device_name()}; building {cfg.name}", file=sys.stderr, flush=True)
    stream = Stream()
    t_build = time.perf_counter()
    scene = sc.DeviceScene(cfg, stream.handle)
    stream.synchronize()
    print(f"[bench] scene ready in {time.perf_counter() - t_build:.1f} s: {scene.grid.num_leaves} leaves, "
          f"majorant {cfg.density_coeff * scene.grid.max_value:.2f}", file=sys.stderr, flush=True)
    pt = PathTracer(scene, stream=stream.handle)
    params = pt.params()

    spp = cfg.spp
    for i in range(args.warmup):
        pt.trace(spp, params)
        stream.synchronize()
        print(f"[bench] warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    pt.stats.fill(0)

    timers = [EventTimer() for _ in range(args.steps)]
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        timers[i].start(stream.handle)
        pt.trace(spp, params)
        timers[i].stop(stream.handle)
    stream.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [t.elapsed_ms() for t in timers]
    c = pt.counters()
    steps = args.steps
    vertices_per_s = c["vertices"] / wall
    avg_launch_s = float(np.mean(kernel_ms)) * 1e-3
    lookups_per_launch = c["lookups"] / steps
    achieved_gbs = lookups_per_launch * BYTES_PER_LOOKUP / avg_launch_s / 1e9
    print(f"[bench] {steps} steps in {wall:.3f} s; kernel avg {np.mean(kernel_ms):.2f} ms; "
          f"{c['paths']} paths, {c['vertices']} vertices, {c['lookups']} lookups", file=sys.stderr, flush=True)
    img = pt.image()
    if not np.all(np.isfinite(img)):
        print("[bench] image has non-finite pixels", file=sys.stderr)
        return 1

    result = {
        "metric": METRIC,
        "value": vertices_per_s,
        "unit": "path vertices/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural fBm volume, Hosek-Wilkie sky)",
        "config": {
            "workload": workload_text(cfg, spp),
            "contrast": cfg.fbm.contrast,
            "tracking": args.tracking,
            "image": [cfg.width, cfg.height],
            "spp_per_step": spp,
        },
        "paths_per_s": c["paths"] / wall,
        "vertices_per_path": c["vertices"] / max(c["paths"], 1),
        "lookups_per_s": c["lookups"] / wall,
        "mean_radiance": [float(v) for v in img.reshape(-1, 3).mean(axis=0)],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved_gbs / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(f"wavefront_kernel_{args.tracking}", cfg.name),
            "kernel": f"wavefront_kernel<PATH_TRACE, {args.tracking.upper()}_MAJORANT>",
            "kernel_ms_avg": float(np.mean(kernel_ms)),
            "algorithmic_bytes_per_launch": lookups_per_launch * BYTES_PER_LOOKUP,
        },
        "cpu_baseline": None,
    }
    result["lookups_per_path"] = c["lookups"] / max(c["paths"], 1)
    if args.legs != "all":
        return finish(args, result, timers, pt, scene, stream, network_only=args.legs == "network")
    waves = (3, 4)
    lat = latency_roofline(scene, params, lookups_per_launch, c["lookups"] / max(c["paths"], 1),
                           c["lookups"] / sum(kernel_ms) * 1e3, c["paths"] / wall, waves=waves)
    result["latency_roofline"] = lat
    print(f"[bench] latency roofline ({lat['tracking']} tracking): march probe {lat['lookups_per_s_3waves']:.3g} "
          f"lookups/s at 3 waves/SIMD ({lat[f'lookups_per_s_{waves[1]}waves']:.3g} at {waves[1]}); trace at "
          f"{lat['frac']:.2f} of it, paths at {lat['paths_frac']:.2f}", file=sys.stderr, flush=True)
    if tracking == _abi.TRACKING_LOCAL_MAJORANT:
        print("[bench] global-majorant leg ...", file=sys.stderr, flush=True)
        glob = global_majorant_leg(scene, pt, spp, max(2, steps // 2), stream)
        glob["local_speedup"] = glob["kernel_ms_avg"] / float(np.mean(kernel_ms))
        result["global_majorant"] = glob
    if args.config == "sparse512":
        print("[bench] saturated-cloud legs ...", file=sys.stderr, flush=True)
        result["saturated_cloud"] = saturated_leg(args, stream, tracking)
        if tracking == _abi.TRACKING_LOCAL_MAJORANT:
            result["saturated_cloud_global_majorant"] = saturated_leg(args, stream, _abi.TRACKING_GLOBAL_MAJORANT)
    if args.nrc_steps > 0:
        print("[bench] NRC network leg ...", file=sys.stderr, flush=True)
        # the network's inputs are drawn over the active leaves of the dense 128^3 cloud
        net_scene = sc.DeviceScene(sc.CONFIGS["dense128"])
        mlp = nrc_mlp_leg(net_scene, args.nrc_steps)
        mlp_hash = nrc_mlp_leg(net_scene, args.nrc_steps, hidden_layers=2, encoding=1)
        mlp_fp8 = nrc_mlp_leg(net_scene, args.nrc_steps, precision=_abi.NRC_BF16_FP8)
        mlp_fp32 = nrc_mlp_leg(net_scene, args.nrc_steps, precision=_abi.NRC_FP32)
        net_scene.close()
        print(f"[bench] NRC network: infer {mlp['infer_ms']:.2f} ms / 2^21, train {mlp['train_ms']:.2f} ms / 2^16; "
              f"e4m3 inference {mlp_fp8['infer_ms']:.2f} ms", file=sys.stderr, flush=True)
        print("[bench] NRC render leg ...", file=sys.stderr, flush=True)
        rend = nrc_render_leg(scene, args.nrc_steps, stream)
        print(f"[bench] NRC render: {rend['ms_per_spp']:.2f} ms/spp", file=sys.stderr, flush=True)
        result["nrc_queries_per_s"] = mlp["queries_per_s"]
        result["nrc"] = {"network": mlp, "network_hash_grid": mlp_hash, "network_fp8": mlp_fp8, "network_fp32": mlp_fp32,
                         "render": rend}
    if args.nrc_steps > 0:
        print("[bench] multi-instance leg ...", file=sys.stderr, flush=True)
        multi = multi_instance_leg(args.nrc_steps, stream)
        result["multi_instance"] = multi
        print(f"[bench] multi-instance: 8 NRC render launches {multi['one_launch_per_instance_ms']:.2f} ms, one "
              f"batched launch {multi['batched_autotuned_ms']:.2f} ms (autotuned: shares by "
              f"{multi['batch_choice']['shares']}, knobs {multi['batch_choice']['knobs']}), "
              f"{multi['batched_by_pixels_ms']:.2f} ms (by pixels)", file=sys.stderr, flush=True)
    if args.rmse_ref_spp > 0:
        print("[bench] equal-RMSE leg ...", file=sys.stderr, flush=True)
        eq = equal_rmse_leg(scene, args.rmse_ref_spp, cfg.spp, stream, crop_ref_spp=args.crop_ref_spp)
        result["equal_rmse"] = eq
        print(f"[bench] equal RMSE: plain rel RMSE {eq['plain_rel_rmse']:.4f} at {cfg.spp} spp; " + "; ".join(
            f"NRC {k} {eq[k]['nrc_rel_rmse']:.4f} -> plain needs {eq[k]['plain_spp_for_equal_rmse']:.0f} spp, "
            f"speedup {eq[k]['speedup'] if eq[k]['speedup'] is None else round(eq[k]['speedup'], 2)}x"
            for k in eq if isinstance(eq[k], dict) and "nrc_rel_rmse" in eq[k]), file=sys.stderr, flush=True)
        print("[bench] equal-RMSE leg on the 128^3 dense cloud ...", file=sys.stderr, flush=True)
        result["equal_rmse_dense128"] = dense128_equal_rmse_leg(args.rmse_ref_spp, stream, crop_ref_spp=args.crop_ref_spp)
    return finish(args, result, timers, pt, scene, stream)


def finish(args, result, timers, pt, scene, stream, network_only: bool = False) -> int:
    """The NRC network legs alone (network_only), the CPU baseline, the JSON line, and the clean-up."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    if network_only and args.nrc_steps > 0:
        net_scene = sc.DeviceScene(sc.CONFIGS["dense128"])
        result["nrc"] = {"network": nrc_mlp_leg(net_scene, args.nrc_steps),
                         "network_hash_grid": nrc_mlp_leg(net_scene, args.nrc_steps, hidden_layers=2, encoding=1),
                         "network_fp8": nrc_mlp_leg(net_scene, args.nrc_steps, precision=_abi.NRC_BF16_FP8)}
        result["nrc_queries_per_s"] = result["nrc"]["network"]["queries_per_s"]
        net_scene.close()
    if args.legs == "dense128" and args.rmse_ref_spp > 0:
        print("[bench] equal-RMSE leg on the 128^3 dense cloud ...", file=sys.stderr, flush=True)
        result["equal_rmse_dense128"] = dense128_equal_rmse_leg(args.rmse_ref_spp, stream, crop_ref_spp=args.crop_ref_spp)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(scene, args.cpu_seconds)
        print(f"[bench] CPU baseline {result['cpu_baseline']['value']:.3g} vertices/s "
              f"({result['cpu_baseline']['seconds']:.1f} s)", file=sys.stderr, flush=True)
    print(json.dumps(result), flush=True)
    for t in timers:
        t.destroy()
    pt.close()
    scene.close()
    stream.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
