"""Post-processing and image files, over the library's post-process kernel and encoders.

``tone_map`` is the batch driver's applyToneMap pass (renderer/gpu_kernels/post_process_kernels.cu:20-40)
fused with saveImage's SDR conversion (common/common_host.cpp:515-570) on the device; ``save_image``
writes PNG, BMP or EXR by extension, as saveImage / saveImageHDR do (common/common_host.cpp:422-513).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _abi
from ._lib import call

# The batch driver's settings (renderer_main.cpp:1490-1495 and post_process_kernels.cu:29,33).
TONE_EXPONENT = 1.0 / 0.6
TONE_BRIGHTNESS = 1.0


def sdr_config(alpha_override: float = 1.0, brightness_scale: float = 1.0, apply_tone_map: bool = False,
               apply_srgb_gamma: bool = True, flip_y: bool = False) -> _abi.SdrConfig:
    """An SDRImageSaverConfig; the defaults are the batch driver's (renderer_main.cpp:1490-1495)."""
    return _abi.SdrConfig(float(alpha_override), float(brightness_scale), int(apply_tone_map),
                          int(apply_srgb_gamma), int(flip_y))


def tone_map(accum_ptr: int, width: int, height: int, samples: int, out_rgba32f_ptr: int | None = None,
             out_rgba8_ptr: int | None = None, sdr: _abi.SdrConfig | None = None, stream=None,
             exponent: float = TONE_EXPONENT, brightness: float = TONE_BRIGHTNESS) -> None:
    """Device pass over an FP64 accumulator (device pointers)."""
    p = _abi.ToneMapParams()
    p.accum = accum_ptr
    p.width, p.height, p.samples = width, height, samples
    p.exponent = np.float32(np.float32(1.0) / np.float32(0.6)) if exponent == TONE_EXPONENT else exponent
    p.brightness = brightness
    p.out_rgba32f = out_rgba32f_ptr
    p.out_rgba8 = out_rgba8_ptr
    p.sdr = sdr if sdr is not None else sdr_config()
    call("rtc8x_tone_map", C.c_void_p(stream), C.byref(p))


def sdr_convert(data: np.ndarray, config: _abi.SdrConfig | None = None) -> np.ndarray:
    """Host SDR conversion of an (H, W, C) float image (C in 1..4) to (H, W) packed RGBA8."""
    img = np.ascontiguousarray(data, dtype=np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    h, w, c = img.shape
    out = np.empty((h, w), dtype=np.uint32)
    call("rtc8x_sdr_convert", img.ctypes.data, w, h, c, C.byref(config or sdr_config()), out.ctypes.data)
    return out


def _encode(name: str, *args) -> bytes:
    size = C.c_size_t(0)
    call(name, *args, None, 0, C.byref(size))
    buf = C.create_string_buffer(size.value)
    call(name, *args, buf, size.value, C.byref(size))
    return buf.raw[: size.value]


def encode_png(rgba8: np.ndarray) -> bytes:
    px = np.ascontiguousarray(rgba8, dtype=np.uint32)
    return _encode("rtc8x_png_encode", px.ctypes.data, px.shape[1], px.shape[0])


def encode_bmp(rgba8: np.ndarray) -> bytes:
    px = np.ascontiguousarray(rgba8, dtype=np.uint32)
    return _encode("rtc8x_bmp_encode", px.ctypes.data, px.shape[1], px.shape[0])


def encode_exr(data: np.ndarray, brightness_scale: float = 1.0, flip_y: bool = False) -> bytes:
    img = np.ascontiguousarray(data, dtype=np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    h, w, c = img.shape
    return _encode("rtc8x_exr_encode", img.ctypes.data, w, h, c, C.c_float(brightness_scale), int(flip_y))


def save_image(path, image: np.ndarray, config: _abi.SdrConfig | None = None, brightness_scale: float = 1.0) -> int:
    """Writes ``image`` by extension: .exr from a float (H, W, C) image; .png / .bmp from packed RGBA8
    (H, W) uint32, or from a float image through the SDR conversion.  Returns the byte count."""
    path = Path(path)
    ext = path.suffix.lower()
    if ext == ".exr":
        data = encode_exr(image, brightness_scale)
    elif ext in (".png", ".bmp"):
        px = image if image.dtype == np.uint32 and image.ndim == 2 else sdr_convert(image, config)
        data = encode_png(px) if ext == ".png" else encode_bmp(px)
    else:
        raise ValueError(f"unsupported image type {ext!r} (png, bmp, exr)")
    path.write_bytes(data)
    return len(data)
