"""Image-error measurement: RMSE against a converged reference and the equal-RMSE wall-time comparison
between the plain tracer and NRC-accelerated rendering.

The reference for a production-size frame is the plain tracer itself at high spp, as two independent
halves A and B (different per-pixel seeds). For an estimate X that is independent of both,
E[(X - A)(X - B)] = E[(X - truth)^2], because the cross terms vanish when A and B are unbiased and
independent. The reference noise therefore drops out of the MSE exactly, in expectation. That the
plain tracer converges to the FP64 CPU oracle's image is checked by the test suite at sizes the oracle
finishes in seconds.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .device import EventTimer
from .nrc import NrcRenderer
from .scene import DeviceScene
from .tracer import PathTracer


def cross_mse(x: np.ndarray, a: np.ndarray, b: np.ndarray) -> float:
    """Unbiased MSE of x against the truth that the independent unbiased references a and b estimate."""
    return float(np.mean((x - a) * (x - b)))


def plain_image(scene: DeviceScene, spp: int, seed: int, stream=None, per_launch: int = 16):
    """Renders `spp` samples per pixel with the plain tracer from the given per-pixel seed; returns the
    mean image (H x W x 3, float64) and the GPU time in seconds."""
    pt = PathTracer(scene, stream=stream)
    pt.reset(seed)
    t = EventTimer()
    t.start(stream)
    done = 0
    while done < spp:
        k = min(per_launch, spp - done)
        pt.trace(k)
        done += k
    t.stop(stream)
    secs = t.elapsed_ms() * 1e-3
    t.destroy()
    img = pt.image()
    pt.close()
    return img, secs


@dataclass
class NrcRun:
    image: np.ndarray
    seconds: float
    spp: int


def nrc_image(scene: DeviceScene, spp: int, seed: int, stream=None, num_hidden_layers: int = 2,
              precision: int = 0, spp_per_frame: int = 1, encoding: int = 1, radiance_scale: float = 0.0,
              ema_decay: float = 0.99, first_frame_iterations: int = 8, train: bool = True,
              train_tracking: int | None = None, loss_scale: float = 0.0) -> NrcRun:
    """NRC-accelerated rendering in the reference's batch schedule (renderer_main.cpp:1554-1644): frame 0
    runs 8 training iterations, later frames one, then `spp_per_frame` samples traced into the cache.
    Wall time includes all training.  train=False renders with the untrained (initial) cache.
    train_tracking: the training launches' tracker (None: the scene's).  loss_scale: the cache's initial
    loss scale (0: 128)."""
    r = NrcRenderer(scene, num_hidden_layers=num_hidden_layers, precision=precision, stream=stream, encoding=encoding,
                    radiance_scale=radiance_scale, ema_decay=ema_decay, first_frame_iterations=first_frame_iterations,
                    train_tracking=train_tracking, loss_scale=loss_scale)
    r.reset(seed)
    t0 = time.perf_counter()
    frames = (spp + spp_per_frame - 1) // spp_per_frame
    for _ in range(frames):
        r.render_frame(spp=spp_per_frame, train_iterations=None if train else 0)
    img = r.image()  # waits for the renderer's stream
    secs = time.perf_counter() - t0
    r.close()
    return NrcRun(img, secs, frames * spp_per_frame)
