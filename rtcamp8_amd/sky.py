"""Hosek-Wilkie sky of the sparse512 workload: loads the per-band model state at turbidity 3 / solar
elevation 30 deg (rtcamp8_amd/data/hosek_T3_e30.npz, produced from the Hosek-Wilkie 2013 model by
tests/golden/make_sky_data.py) and bakes the lat-long environment texture on the GPU
(rtc8x_sky_bake, the role of generateArHosekSkyEnvironmentalTexture)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _abi
from ._lib import call
from .device import DeviceBuffer

DATA = Path(__file__).resolve().parent / "data" / "hosek_T3_e30.npz"


class HosekSky:
    """Device-resident model tables plus the rtc8x_hosek_sky block that points at them."""

    def __init__(self, path: Path = DATA):
        d = np.load(path)
        self.bands = DeviceBuffer.from_array(np.ascontiguousarray(d["bands"], dtype=np.float32))
        self.solar = DeviceBuffer.from_array(np.ascontiguousarray(d["solar_rows"], dtype=np.float32))
        self.limb = DeviceBuffer.from_array(np.ascontiguousarray(d["limb_darkening"], dtype=np.float32))
        s = _abi.HosekSky()
        s.bands, s.solar_rows, s.limb_darkening = self.bands.ptr, self.solar.ptr, self.limb.ptr
        s.solar_turbidity_base = int(d["solar_turbidity_base"])
        s.cmf_x[:] = d["cmf_x"].tolist()
        s.cmf_y[:] = d["cmf_y"].tolist()
        s.cmf_z[:] = d["cmf_z"].tolist()
        s.center_wavelengths[:] = d["center_wavelengths"].tolist()
        s.integral_cmf = float(d["integral_cmf"])
        self.struct = s
        self.turbidity = float(d["turbidity"])
        self.elevation = float(d["elevation"])

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
        call("rtc8x_sky_eval", C.c_void_p(stream), C.byref(self.struct), sd, n, C.c_void_p(dp.ptr),
             C.c_void_p(dt.ptr), C.c_void_p(out.ptr))
        return out.download(np.float32, 3 * n).reshape(n, 3)

    def close(self) -> None:
        for b in (self.bands, self.solar, self.limb):
            b.free()


def bake_hosek_sky(cfg, sun_dir, stream=None) -> DeviceBuffer:
    sky = HosekSky()
    assert abs(sky.turbidity - cfg.turbidity) < 1e-6, "the shipped sky data is for turbidity 3"
    tex = sky.bake(sun_dir, cfg.sky_width, cfg.sky_height, stream)
    from .device import synchronize

    synchronize()
    sky.close()
    return tex
