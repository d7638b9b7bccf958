"""The measured workloads at their stated shapes.

* Dense 128^3 (sigma_t max 20, albedo 0.99, HG g = 0.8, constant white sky + sun, 256 x 256, 4 spp,
  <= 16 bounces): the GPU tracer against the FP32 oracle bit for bit, both trackers.
* The same scene with the NRC integrated, online-trained in the reference's frame schedule
  (renderer/renderer_main.cpp:1554-1644), BF16 and FP32 networks: relative RMSE of the NRC image and of
  the plain tracer at equal samples against the FP64 oracle, on a reduced image.  The numbers are
  printed and pinned (what the trained cache gains, and the bias it costs), not hidden.
* Eight uneven instances (64^3 to 1024^3 fBm, sigma_t max 5 to 200) in one persistent launch with
  autotuned scheduling knobs and shares, each with its own NRC in BF16 training and E4M3 inference, at
  reduced images: every instance's trace outputs equal its own launch's, and those single launches equal
  the FP32 oracle.
"""
import ctypes as C
import time

import numpy as np
import pytest

from oracle_lib import (HostNrc, download_env, download_grid, nthreads, oracle, oracle_seed, ptr,
                        trace_parity_case)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tracking", [0, 1])
def test_dense128_stated_shape_bit_exact(tracking):
    """The full dense 128^3 workload: 256 x 256, 4 spp in one launch, against the FP32 oracle."""
    print(f"[configs] dense128 256x256x4, tracking {tracking}", flush=True)
    res = trace_parity_case("dense128", width=256, height=256, spp=4, tracking=tracking)
    assert res["rng_equal"]
    assert res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))
    assert res["paths"] == 256 * 256 * 4
    assert res["lookups"] == int(res["cpu_stats"][2])


# ------------------------------------------------------------------ NRC on the dense 128^3 cloud
W4, H4 = 96, 96


def _fp64_image(ds, spp: int, seed: int) -> np.ndarray:
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    cfg = ds.cfg.with_(tracking=_abi.TRACKING_GLOBAL_MAJORANT)
    hg = download_grid(ds.grid)
    he, keep = download_env(ds.env)
    p = sc.fill_params(cfg, hg.grid, he, ds.camera)
    rng = oracle_seed(cfg.width, cfg.height, seed)
    accum = np.zeros(3 * cfg.width * cfg.height)
    p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), ptr(accum), None, None
    assert oracle().oracle_trace_f64(C.byref(p), _abi.ENTRY_PATH_TRACE, spp, 0, cfg.height, nthreads()) == 0
    return accum.reshape(cfg.height, cfg.width, 3) / spp


def test_dense128_nrc_against_plain_and_the_fp64_oracle():
    """16 spp each; the judge is the FP64 oracle as two independent 256-spp halves A and B, so that
    mean((X - A)(X - B)) is X's MSE against the converged image with the oracle's noise removed.  The NRC
    trains with the default initial loss scale (128, every step applied, as in the reference) and with
    2^22 (the fresh network's first steps, whose FP16 loss gradients overflow at that scale, skipped).

    Pinned (measured on MI355X at 96 x 96, triangle-wave encoding: plain 0.1015 relative RMSE, all of it
    noise (0.1023); initial scale 128: NRC BF16 0.1455 with noise 0.0691, FP32 0.1496 with noise 0.0699;
    initial scale 2^22: BF16 0.0570 with noise 0.0469, FP32 0.0574 with noise 0.0504; untrained cache
    0.454):
    * the plain tracer is unbiased: its error is its noise, about 0.1 relative RMSE at 16 spp here;
    * the online-trained cache takes noise out (two NRC renders with different seeds differ less than two
      plain renders do: the variance it removes from every path suffix);
    * what it costs is bias, the NRC's relative RMSE against the oracle being its noise plus the cache's
      error.  With every step applied the NRC's error is above the plain tracer's at 16 spp on this small
      scene; with the early steps skipped it is about half of the plain tracer's.  BF16 and FP32 networks
      land within 15 % of each other at either scale;
    * training is what gains: the untrained cache's error is several times the trained ones'."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.convergence import cross_mse, nrc_image, plain_image

    cfg = sc.CONFIGS["dense128"].with_(width=W4, height=H4)
    ds = sc.DeviceScene(cfg)
    try:
        t = time.perf_counter()
        a = _fp64_image(ds, 256, 0xA11CE)
        b = _fp64_image(ds, 256, 0xB0B)
        print(f"[configs] FP64 oracle 2 x 256 spp at {W4}x{H4} in {time.perf_counter() - t:.1f} s", flush=True)
        mean = float((a + b).mean() / 2)
        spp = 16
        x1, tx = plain_image(ds, spp, seed=0xC0FFEE)
        x2, _ = plain_image(ds, spp, seed=0xFACADE)
        plain_rel = np.sqrt(max(cross_mse(x1, a, b), 0.0)) / mean
        plain_noise = np.sqrt(np.mean((x1 - x2) ** 2) / 2) / mean
        res = {}
        for scale in (0.0, 2.0 ** 22):
            for name, prec in (("bf16", _abi.NRC_BF16), ("fp32", _abi.NRC_FP32)):
                runs = [nrc_image(ds, spp, seed=s, precision=prec, spp_per_frame=4, encoding=_abi.NRC_TRIANGLE_WAVE,
                                  radiance_scale=10.0, ema_decay=0.95, loss_scale=scale) for s in (0xD00D, 0xBEEF)]
                rel = np.sqrt(max(cross_mse(runs[0].image, a, b), 0.0)) / mean
                noise = np.sqrt(np.mean((runs[0].image - runs[1].image) ** 2) / 2) / mean
                res[scale, name] = (rel, noise, runs[0].seconds)
        untrained = nrc_image(ds, spp, seed=0xD00D, precision=_abi.NRC_BF16, spp_per_frame=4,
                              encoding=_abi.NRC_TRIANGLE_WAVE, radiance_scale=10.0, ema_decay=0.95, train=False)
        untrained_rel = np.sqrt(max(cross_mse(untrained.image, a, b), 0.0)) / mean
    finally:
        ds.close()
    print(f"[configs] dense128 NRC vs plain at {W4}x{H4}, {spp} spp, against the FP64 oracle: plain rel RMSE "
          f"{plain_rel:.4f} (noise {plain_noise:.4f}, {tx * 1e3:.2f} ms); " + "; ".join(
              f"NRC {name} at initial scale {scale or 128:g}: {r[0]:.4f} (noise {r[1]:.4f}, {r[2] * 1e3:.1f} ms)"
              for (scale, name), r in res.items()) + f"; untrained cache {untrained_rel:.4f}", flush=True)
    assert 0.08 < plain_rel < 0.13 and abs(plain_noise / plain_rel - 1) < 0.15
    for (scale, name), (rel, noise, _) in res.items():
        if scale == 0.0:  # every step applied: bias included, above the plain tracer's error
            assert np.isfinite(rel) and 0.1 < rel < 0.2, (scale, name, rel)
            assert noise < 0.8 * plain_noise, (scale, name, noise, plain_noise)
        else:  # the early steps skipped: below the plain tracer's error
            assert np.isfinite(rel) and 0.03 < rel < 0.08, (scale, name, rel)
            assert rel < 0.8 * plain_rel, (scale, name, rel, plain_rel)
            assert noise < 0.6 * plain_noise, (scale, name, noise, plain_noise)
    for scale in (0.0, 2.0 ** 22):
        assert abs(res[scale, "bf16"][0] / res[scale, "fp32"][0] - 1) < 0.15, scale
    assert untrained_rel > 2.0 * max(r[0] for r in res.values())


# ------------------------------------------------------------------ eight instances in one launch
INSTANCES = [  # fBm resolution, sparse, sigma_t max, image (reduced from 256^2 .. 1920 x 1080)
    (64, False, 5.0, 64, 64), (96, False, 10.0, 80, 60), (128, False, 20.0, 96, 96), (192, True, 40.0, 80, 45),
    (256, True, 60.0, 100, 75), (384, True, 100.0, 96, 54), (512, True, 150.0, 128, 72), (1024, True, 200.0, 160, 90),
]


def test_eight_instances_fp8_inference_autotuned_against_single_launches_and_the_oracle():
    from rtcamp8_amd import _abi, batch
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.nrc import NrcRenderer

    base = sc.CONFIGS["sparse512"]
    scenes, rs = [], []
    for res, sparse, sigma, w, h in INSTANCES:
        print(f"[configs] instance {res}^3 sigma {sigma} {w}x{h}", flush=True)
        fbm = sc.FbmConfig(resolution=res, seed=0x5EED0100 + res, sparse=sparse, active_leaf_fraction=0.25)
        cfg = base.with_(name=f"fbm{res}", width=w, height=h, density_coeff=sigma, fbm=fbm, max_path_length=64,
                         sky_width=256, sky_height=128)
        scenes.append(sc.DeviceScene(cfg))
        rs.append(NrcRenderer(scenes[-1], num_hidden_layers=2, precision=_abi.NRC_BF16_FP8, samples_per_launch=1,
                              radiance_scale=10.0, ema_decay=0.95, first_frame_iterations=2))
    try:
        for r in rs:  # the first frame trains the caches (BF16), inference runs in E4M3
            r.render_frame(spp=1)
        keys = ("inference_queries", "inference_terminals", "per_frame_contribution")

        def params(r):
            p = r.params()
            p.nrc.num_samples = 1
            return p

        def outputs(r):
            return [getattr(r.buffers, k).download(np.uint32, getattr(r.buffers, k).nbytes // 4) for k in keys]

        shares, knobs = batch.autotune(
            lambda i, ev, shape: call("rtc8x_trace", None, C.byref(batch.apply_knobs(params(rs[i]), (ev, shape))),
                                      _abi.ENTRY_PATH_TRACE_WITH_NRC, 1), len(rs), repeats=1)
        assert len(shares) == 8 and all(s > 0 for s in shares)
        print(f"[configs] autotuned (event eighths, tile shape) per instance: {knobs}", flush=True)
        for r, k in zip(rs, knobs):
            r.knobs = k
        rng0 = [r.rng.download(np.uint64, r.width * r.height) for r in rs]
        single, before = [], []
        for r, g in zip(rs, rng0):  # each instance's own launch, on cleared output buffers
            r.rng.upload(g)
            for k in keys:
                getattr(r.buffers, k).fill(0)
            before.append(HostNrc(r.buffers))
            call("rtc8x_trace", None, C.byref(params(r)), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1)
            single.append((outputs(r), r.rng.download(np.uint64, r.width * r.height)))
        for r, g in zip(rs, rng0):
            r.rng.upload(g)
            for k in keys:
                getattr(r.buffers, k).fill(0)
        batch._launch([params(r) for r in rs], [1] * 8, _abi.ENTRY_PATH_TRACE_WITH_NRC, shares, None)
        for i, r in enumerate(rs):
            for k, want, got in zip(keys, single[i][0], outputs(r)):
                assert np.array_equal(want, got), (i, k)
            assert np.array_equal(single[i][1], r.rng.download(np.uint64, r.width * r.height)), i
        # the single launches against the FP32 oracle on the same inputs
        for i, (r, g) in enumerate(zip(rs, rng0)):
            print(f"[configs] instance {i}: oracle", flush=True)
            cfg = r.scene.cfg
            hg = download_grid(r.scene.grid)
            he, keep = download_env(r.scene.env)
            p = sc.fill_params(cfg, hg.grid, he, r.scene.camera)
            host = before[i]
            host.fill(p.nrc)
            dp = params(r)
            for f in ("nrc_buffer_index", "offset_to_select_unbiased_path", "radiance_scale", "path_termination_factor",
                      "num_samples"):
                setattr(p.nrc, f, getattr(dp.nrc, f))
            rng = g.copy()
            p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), None, None, None
            assert oracle().oracle_trace_f32(C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1, 0, cfg.height,
                                             nthreads()) == 0
            n = cfg.width * cfg.height
            assert np.array_equal(rng, single[i][1]), i
            term = host.a["inference_terminals"].reshape(-1, 4)[:n]
            assert np.array_equal(term, single[i][0][1].reshape(-1, 4)[:n]), i
            has = (term[:, 3] & 1).astype(bool)
            want_q = single[i][0][0].reshape(-1, 14)[:n][has]
            assert np.array_equal(host.a["inference_queries"].reshape(-1, 14)[:n][has].view(np.uint32), want_q), i
            assert np.array_equal(host.a["per_frame_contribution"].view(np.uint32)[:3 * n], single[i][0][2][:3 * n]), i
        batch.render_nrc(rs, [1] * 8, shares)
        for r in rs:
            img = r.image()
            assert np.all(np.isfinite(img)) and img.mean() > 0
    finally:
        for r in rs:
            r.close()
        for s in scenes:
            s.close()
