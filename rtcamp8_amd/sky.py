"""Hosek-Wilkie skies for any sun: the spectral model's tables (rtcamp8_amd/data/hosek_model.bin,
written from the Hosek-Wilkie 2013 model by tests/golden/make_sky_model.py) give the per-band state for
a solar elevation and turbidity (rtc8x_hosek_model_state, the host part of
AnalyticSkyEnvironment::computeDistribution, renderer/renderer_scene.cpp:555-602), and the lat-long
environment texture is baked on the GPU (rtc8x_sky_bake, the role of
generateArHosekSkyEnvironmentalTexture)."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import _abi
from ._lib import call
from .device import DeviceBuffer, wait_stream

MODEL = Path(__file__).resolve().parent / "data" / "hosek_model.bin"


class HosekModel:
    """The model tables, loaded once; ``state`` re-derives the sky block for a sun."""

    def __init__(self, path: Path = MODEL):
        blob = Path(path).read_bytes()
        self.handle = C.c_void_p()
        call("rtc8x_hosek_model_create", blob, len(blob), C.byref(self.handle))

    def bands(self, solar_elevation: float, turbidity: float) -> np.ndarray:
        """The 16 per-band states as a [16, 136] float32 array (host)."""
        out = np.zeros((_abi.SKY_BANDS, 136), np.float32)
        call("rtc8x_hosek_bands", self.handle, float(solar_elevation), float(turbidity), out.ctypes.data)
        return out

    def state(self, solar_elevation: float, turbidity: float, stream=None) -> _abi.HosekSky:
        sky = _abi.HosekSky()
        call("rtc8x_hosek_model_state", self.handle, float(solar_elevation), float(turbidity), C.byref(sky),
             C.c_void_p(stream))
        return sky

    def close(self) -> None:
        if self.handle:
            call("rtc8x_hosek_model_destroy", self.handle)
            self.handle = C.c_void_p()


def solar_elevation_of(sun_dir) -> float:
    """0.5 pi - acos(clamp(sun.y)) in float32, as computeDistribution derives it (renderer_scene.cpp:570)."""
    y = np.float32(min(max(float(sun_dir[1]), -1.0), 1.0))
    return float(np.float32(0.5) * np.float32(math.pi) - np.float32(math.acos(float(y))))


class HosekSky:
    """A sky for one sun: the model's state for (solar elevation, turbidity) on the device."""

    def __init__(self, solar_elevation: float = math.radians(30.0), turbidity: float = 3.0, model: HosekModel | None = None):
        self.own_model = model is None
        self.model = model or HosekModel()
        self.elevation = float(np.float32(solar_elevation))
        self.turbidity = float(np.float32(turbidity))
        self.struct = self.model.state(self.elevation, self.turbidity)

    def bake(self, sun_dir, width: int, height: int, stream=None) -> DeviceBuffer:
        tex = DeviceBuffer(4 * 4 * width * height)
        sd = (C.c_float * 3)(*[float(v) for v in sun_dir])
        call("rtc8x_sky_bake", C.c_void_p(stream), C.byref(self.struct), sd, width, height, C.c_void_p(tex.ptr))
        return tex

    def eval(self, sun_dir, phi: np.ndarray, theta: np.ndarray, stream=None) -> np.ndarray:
        n = len(phi)
        dp = DeviceBuffer.from_array(np.ascontiguousarray(phi, dtype=np.float32))
        dt = DeviceBuffer.from_array(np.ascontiguousarray(theta, dtype=np.float32))
        out = DeviceBuffer(12 * n)
        sd = (C.c_float * 3)(*[float(v) for v in sun_dir])
        try:
            call("rtc8x_sky_eval", C.c_void_p(stream), C.byref(self.struct), sd, n, C.c_void_p(dp.ptr),
                 C.c_void_p(dt.ptr), C.c_void_p(out.ptr))
            wait_stream(stream)
            return out.download(np.float32, 3 * n).reshape(n, 3)
        finally:
            for b in (dp, dt, out):
                b.free()

    def close(self) -> None:
        if self.own_model:
            self.model.close()


def bake_hosek_sky(cfg, sun_dir, stream=None) -> DeviceBuffer:
    """The workload's sky texture (cfg.sky_width x cfg.sky_height) for its sun and turbidity."""
    from .device import synchronize

    sky = HosekSky(solar_elevation_of(sun_dir), cfg.turbidity)
    tex = sky.bake(sun_dir, cfg.sky_width, cfg.sky_height, stream)
    synchronize()
    sky.close()
    return tex
