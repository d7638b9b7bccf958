"""Benchmark of the plain delta-tracking volumetric path tracer on one MI355X.

Workload (sparse512): sparse 512^3 NanoVDB-layout fBm cloud (25 % of 8^3 leaves active, sigma_t max 50),
albedo 0.999, Henyey-Greenstein g = 0.85, Hosek-Wilkie sky (turbidity 3, sun elevation 30 deg) sampled
with alias tables, 1920x1080, up to 256 bounces, Russian roulette after bounce 8.  One step = one sample
per pixel over the full image (one launch); the default K = 16 steps is the workload's 16 spp.

Prints one JSON line: value = path vertices per second (rays traced through the medium per second,
all inputs resident in HBM), with the trace kernel's roofline (HIP-event kernel time, 32 algorithmic
bytes per trilinear density lookup) and the CPU oracle timed on the host's cores as a baseline.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="sparse512")
    ap.add_argument("--tracking", choices=["local", "global"], default="global")
    ap.add_argument("--cpu-rows", type=int, default=6, help="image rows of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if there is one."""
    f = ROOT / "profiles" / f"{kernel}_{workload}_hbm.json"
    if not f.exists():
        return None
    try:
        return float(json.loads(f.read_text())["hbm_bytes_per_launch"])
    except Exception:
        return None


def cpu_baseline(scene, args, rows: int):
    """The CPU oracle (FP32 restatement of the same algorithm) on `rows` image rows, 1 spp, all cores."""
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
    y0 = h // 2 - rows // 2
    t0 = time.perf_counter()
    rc = oracle().oracle_trace_f32(C.byref(p), _abi.ENTRY_PATH_TRACE, 1, y0, y0 + rows, threads)
    dt = time.perf_counter() - t0
    assert rc == 0
    vertices = int(stats[_abi.STAT_VERTICES])
    return {"value": vertices / dt, "unit": "path vertices/s", "cores": threads, "kind": "port",
            "sample": f"{rows} rows x {w} px at 1 spp through the image centre ({rows * w} paths, {vertices} vertices, "
                      f"{dt:.1f} s), FP32 oracle",
            "seconds": dt}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); run it with --gpus 1", file=sys.stderr)
        return 2
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

    for i in range(args.warmup):
        pt.trace(1, params)
        stream.synchronize()
        print(f"[bench] warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    pt.stats.fill(0)

    timers = [EventTimer() for _ in range(args.steps)]
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        timers[i].start(stream.handle)
        pt.trace(1, params)
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
            "workload": f"{cfg.name}: sparse {cfg.fbm.resolution}^3 fBm cloud, 25% of 8^3 leaves, sigma_t max "
                        f"{cfg.density_coeff:g}, albedo {cfg.albedo}, HG g={cfg.forwardness}, Hosek-Wilkie T=3 sun 30deg, "
                        f"alias-table env sampling, {cfg.width}x{cfg.height}, {steps} spp, <= {cfg.max_path_length} bounces, "
                        f"RR after {cfg.rr_after_path_length}",
            "tracking": args.tracking,
            "image": [cfg.width, cfg.height],
            "spp": steps,
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
            "traffic": traffic_from_profiles("trace_kernel", cfg.name),
            "kernel": "trace_kernel<PATH_TRACE>",
            "kernel_ms_avg": float(np.mean(kernel_ms)),
            "algorithmic_bytes_per_launch": lookups_per_launch * BYTES_PER_LOOKUP,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(scene, args, args.cpu_rows)
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
