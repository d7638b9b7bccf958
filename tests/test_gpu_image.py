"""The post-process kernel on the device: the tone-mapped frame and its packed SDR pixels equal the
oracle's FP32 restatement bit for bit, and the double-precision evaluation to within rounding."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flip,tone", [(False, False), (True, True)])
def test_tone_map_matches_oracle(flip, tone):
    from rtcamp8_amd import image
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.device import DeviceBuffer
    from rtcamp8_amd.tracer import PathTracer
    from oracle_lib import oracle

    w, h, spp = 67, 45, 4
    cfg = sc.CONFIGS["dense128"].with_(width=w, height=h, spp=spp, fbm=sc.FbmConfig(resolution=64, seed=9))
    ds = sc.DeviceScene(cfg)
    pt = PathTracer(ds)
    print(f"[image] trace {w}x{h} flip={flip} tone={tone}", flush=True)
    pt.trace(spp)
    out32, out8 = DeviceBuffer(16 * w * h), DeviceBuffer(4 * w * h)
    sdr = image.sdr_config(flip_y=flip, apply_tone_map=tone, brightness_scale=1.7)
    try:
        image.tone_map(pt.accum.ptr, w, h, spp, out32.ptr, out8.ptr, sdr)
        got32 = out32.download(np.float32, 4 * w * h)
        got8 = out8.download(np.uint32, w * h)
        accum = pt.accum.download(np.float64, 3 * w * h)
    finally:
        for b in (out32, out8):
            b.free()
        pt.close()
        ds.close()
    exponent = np.float32(1.0) / np.float32(0.6)
    for f64 in (0, 1):
        want32, want8 = np.zeros(4 * w * h, np.float32), np.zeros(w * h, np.uint32)
        oracle().oracle_tone_map(accum.ctypes.data, w, h, spp, float(exponent), 1.0, C.byref(sdr), want32.ctypes.data,
                                 want8.ctypes.data, f64)
        if f64 == 0:
            assert np.array_equal(got32.view(np.uint32), want32.view(np.uint32))
            assert np.array_equal(got8, want8)
        else:
            assert np.allclose(got32, want32, rtol=2e-5, atol=1e-7)
            d = np.abs(got8.view(np.uint8).astype(np.int32) - want8.view(np.uint8).astype(np.int32))
            assert d.max() <= 1 and np.mean(d > 0) < 2e-3
    assert np.all(got32.reshape(-1, 4)[:, 3] == 1.0) and np.all(got8 >> 24 == 255)
