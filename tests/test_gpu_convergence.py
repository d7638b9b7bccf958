"""Converged images: the gfx950 tracer against the FP64 CPU oracle on independent random streams
(statistical equivalence, for both trackers), RMSE proportional to 1/sqrt(spp), and a white furnace
whose expected image is known in closed form."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import download_env, download_grid, nthreads, oracle, oracle_seed, ptr

pytestmark = pytest.mark.gpu

W, H = 64, 48


def _cfg(**kw):
    from rtcamp8_amd import scene as sc

    base = sc.CONFIGS["dense128"].with_(width=W, height=H, fbm=sc.FbmConfig(resolution=64, seed=0x5EED0001))
    return base.with_(**kw)


def _gpu(ds, spp, seed, tracking=None):
    from rtcamp8_amd.convergence import plain_image

    if tracking is not None:
        ds.cfg = ds.cfg.with_(tracking=tracking)
    img, _ = plain_image(ds, spp, seed)
    return img


def _fp64(ds, spp, seed):
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc

    cfg = ds.cfg.with_(tracking=0)
    hg = download_grid(ds.grid)
    he, keep = download_env(ds.env)
    p = sc.fill_params(cfg, hg.grid, he, ds.camera)
    rng = oracle_seed(W, H, seed)
    accum = np.zeros(3 * W * H)
    p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), ptr(accum), None, None
    assert oracle().oracle_trace_f64(C.byref(p), _abi.ENTRY_PATH_TRACE, spp, 0, H, nthreads()) == 0
    return accum.reshape(H, W, 3) / spp


@pytest.mark.parametrize("tracking", [0, 1])
def test_converges_to_fp64_oracle(tracking):
    """GPU (seed 1) vs FP64 oracle (seed 2) differ exactly as much as two GPU runs (seeds 1, 3) do:
    E|A - B|^2 = 2 sigma^2 / spp for unbiased estimates of the same image."""
    from rtcamp8_amd import scene as sc

    spp = 256
    ds = sc.DeviceScene(_cfg())
    try:
        a = _gpu(ds, spp, 1001, tracking)
        c = _gpu(ds, spp, 3003, 0)
        b = _fp64(ds, spp, 2002)
    finally:
        ds.close()
    m_ab = float(np.mean((a - b) ** 2))
    m_ac = float(np.mean((a - c) ** 2))
    ratio = m_ab / m_ac
    assert 0.8 < ratio < 1.25, ratio
    # no bias: the image means agree within 4 standard errors of the pixel differences
    diff = (a - b).mean()
    se = np.sqrt(m_ac / a.size)
    assert abs(diff) < 4 * se + 1e-12, (diff, se)
    # the stated error: relative RMSE of one 256-spp GPU image against the 256-spp FP64 oracle image
    rel = np.sqrt(m_ab) / b.mean()
    assert rel < 0.05, rel


def test_rmse_scales_as_inverse_sqrt_spp():
    from rtcamp8_amd import scene as sc

    ds = sc.DeviceScene(_cfg())
    try:
        ref = _gpu(ds, 4096, 77)
        rmse = [np.sqrt(np.mean((_gpu(ds, n, 88) - ref) ** 2)) for n in (16, 64, 256)]
    finally:
        ds.close()
    # sigma^2 (1/n + 1/4096): the ratio for 4x the samples is 2 (1.96 for 64 -> 256)
    r1, r2 = rmse[0] / rmse[1], rmse[1] / rmse[2]
    assert 1.8 < r1 < 2.2 and 1.75 < r2 < 2.15, (r1, r2)


def test_white_furnace():
    """Albedo 1, constant unit sky, no sun, unbounded paths: every camera path carries the pinhole weight
    az^3 * sensor_area (optix_pathtracing_kernels.cu:43-55), so a pixel's expectation is that weight
    averaged over the pixel, whatever the medium does."""
    from rtcamp8_amd import scene as sc

    cfg = _cfg(albedo=1.0, env_kind="constant", env_radiance=(1.0, 1.0, 1.0), sun_enabled=False,
               density_coeff=5.0, max_path_length=0, rr_after_path_length=0)
    ds = sc.DeviceScene(cfg)
    try:
        img = _gpu(ds, 512, 4242)
        vh, vw = ds.camera.vh, ds.camera.vw
    finally:
        ds.close()
    s = 32
    sub = (np.arange(s) + 0.5) / s
    expect = np.zeros((H, W))
    for y in range(H):
        py = (y + sub[:, None]) / H
        for x in range(W):
            px = (x + sub[None, :]) / W
            lx, ly = vw * (-0.5 + px), vh * (0.5 - py)
            az = 1.0 / np.sqrt(1.0 + lx * lx + ly * ly)
            expect[y, x] = np.mean(az ** 3) * vw * vh
    ratio = img / expect[..., None]
    se = ratio.std() / np.sqrt(ratio.size)
    assert abs(ratio.mean() - 1.0) < 4 * se + 1e-4, (ratio.mean(), se)
    assert np.all(np.isfinite(img)) and np.all(img >= 0)
