"""GPU Hosek-Wilkie bake vs FP64 ground truth from the vendored model (tests/golden/hosek_golden.npz),
and the environment samplers built from the baked texture."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden" / "hosek_golden.npz"


def test_sky_eval_matches_fp64_model():
    from rtcamp8_amd.sky import HosekSky

    g = np.load(GOLDEN)
    sky = HosekSky(math.radians(30.0), 3.0)
    got = sky.eval(g["sun"], g["phi"], g["theta"]).astype(np.float64)
    ref = g["rgb"]
    sky_part, disc = slice(0, 1152), slice(1152, None)
    rel = np.abs(got[sky_part] - ref[sky_part]) / np.maximum(np.abs(ref[sky_part]), 1e-3)
    # FP32 evaluation of the model vs FP64 away from the sun: 2e-4 relative per channel
    assert np.max(rel) < 2e-4, (np.max(rel), np.argmax(rel) // 3)
    # Around the solar disc limb darkening has an infinite slope at the rim, so FP32 direction
    # rounding moves rim samples: bound their error by 2e-3 of the disc's peak, and check samples well
    # inside the disc to 1e-3 relative.  There d(limb cosine)/d(gamma) is below ~220 per radian and an
    # FP32 view direction is good to a few 1e-7 rad, so the FP32-vs-FP64 difference stays below ~5e-4.
    peak = ref[disc].max()
    assert np.max(np.abs(got[disc] - ref[disc])) < 2e-3 * peak
    phi, theta = g["phi"][disc].astype(np.float64), g["theta"][disc].astype(np.float64)
    view = np.stack([-np.sin(phi) * np.sin(theta), np.cos(theta), np.cos(phi) * np.sin(theta)], axis=1)
    gamma = np.arccos(np.clip(view @ g["sun"].astype(np.float64), -1, 1))
    radius = 0.51 * math.pi / 180 / 2
    limb_cos = np.sqrt(np.clip(1 - (np.sin(gamma) / math.sin(radius)) ** 2, 0, None))
    inner = limb_cos > 0.7
    assert inner.sum() >= 3
    rel_in = np.abs(got[disc][inner] - ref[disc][inner]) / ref[disc][inner]
    assert np.max(rel_in) < 1e-3, np.max(rel_in)
    # the solar disc is present and orders of magnitude brighter than the sky
    assert peak > 1000 * np.median(ref[sky_part])
    sky.close()


def test_bake_matches_eval_at_texel_centres():
    from rtcamp8_amd.sky import HosekSky

    g = np.load(GOLDEN)
    sky = HosekSky(math.radians(30.0), 3.0)
    w, h = 64, 32
    tex = sky.bake(g["sun"], w, h).download(np.float32, 4 * w * h).reshape(h, w, 4)
    y, x = np.mgrid[0:h, 0:w]
    theta = (np.float32(math.pi) * (y.astype(np.float32) + np.float32(0.5)) / np.float32(h)).ravel()
    phi = (np.float32(2 * math.pi) * (x.astype(np.float32) + np.float32(0.5)) / np.float32(w)).ravel()
    ev = sky.eval(g["sun"], phi, theta).reshape(h, w, 3)
    assert np.allclose(tex[..., :3], ev, rtol=1e-6, atol=0)
    assert np.all(tex[..., 3] == 1.0)
    sky.close()


def test_second_sun_matches_fp64_model():
    """A sky re-derived from the model tables for another sun (elevation 70 deg, turbidity 6.5), as the
    batch driver does per frame, against the vendored FP64 model (tests/golden/hosek_states.npz)."""
    from rtcamp8_amd.sky import HosekSky

    g = np.load(Path(__file__).resolve().parent / "golden" / "hosek_states.npz")
    sky = HosekSky(float(g["sun2_elevation"]), float(g["sun2_turbidity"]))
    got = sky.eval(g["sun2"], g["phi2"], g["theta2"]).astype(np.float64)
    ref = g["rgb2"]
    sky_part, disc = slice(0, 1152), slice(1152 + 64, None)
    rel = np.abs(got[sky_part] - ref[sky_part]) / np.maximum(np.abs(ref[sky_part]), 1e-3)
    assert np.max(rel) < 2e-4, np.max(rel)
    peak = ref[disc].max()
    assert np.max(np.abs(got[disc] - ref[disc])) < 2e-3 * peak
    assert peak > 1000 * np.median(ref[sky_part])
    sky.close()
