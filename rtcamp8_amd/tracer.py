"""The plain delta-tracking path tracer driven through the C ABI: the `renderer == 0` branch of the
reference's frame loop (renderer/renderer_main.cpp:1639-1643), one launch of __raygen__pathTrace per
batch of samples, accumulated in FP64."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._lib import call, lib
from .device import DeviceBuffer, wait_stream
from .scene import DeviceScene


class PathTracer:
    def __init__(self, scene: DeviceScene, stream=None, track_vertices: bool = False, schedule: bool = True):
        self.scene = scene
        self.stream = stream
        cfg = scene.cfg
        n = cfg.width * cfg.height
        self.rng = DeviceBuffer(8 * n)
        self.accum = DeviceBuffer(24 * n)
        self.vertices = DeviceBuffer(4 * n) if track_vertices else None
        self.stats = DeviceBuffer(8 * _abi.STAT_COUNT)
        # cost-ordered scheduling: each launch records per-pixel lookups, the next starts the costliest
        # 8 x 8 tiles first
        self.schedule = schedule
        self.cost = DeviceBuffer.zeros(4 * n) if schedule else None
        self.order = DeviceBuffer(4 * n) if schedule else None
        self.order_scratch = (DeviceBuffer(lib().rtc8x_order_tiles_scratch_bytes(cfg.width, cfg.height))
                              if schedule else None)
        self.have_cost = False
        self.samples = 0
        self.knobs = (0, 0)  # (event_eighths, tile_shape) of the trace kernel's schedule; 0 = defaults
        self.reset()

    def reset(self, seed: int | None = None) -> None:
        cfg = self.scene.cfg
        call("rtc8x_seed_rng", C.c_void_p(self.rng.ptr), cfg.width, cfg.height,
             C.c_uint64(cfg.rng_seed if seed is None else seed))
        self.accum.fill(0)
        self.stats.fill(0)
        if self.vertices is not None:
            self.vertices.fill(0)
        self.samples = 0

    def params(self) -> _abi.FrameParams:
        p = self.scene.params()
        p.rng_states = self.rng.ptr
        p.accum = self.accum.ptr
        p.path_vertices = self.vertices.ptr if self.vertices is not None else None
        p.stats = self.stats.ptr
        p.event_eighths, p.tile_shape = self.knobs
        return p

    def _schedule(self, p: _abi.FrameParams) -> None:
        if not self.schedule:
            return
        cfg = self.scene.cfg
        if self.have_cost:
            call("rtc8x_order_by_tile_cost", C.c_void_p(self.stream), C.c_void_p(self.cost.ptr), cfg.width, cfg.height,
                 C.c_void_p(self.order.ptr), C.c_void_p(self.order_scratch.ptr))
            p.work_order = self.order.ptr
        else:
            p.work_order = None
        p.pixel_cost = self.cost.ptr
        self.have_cost = True

    def trace(self, samples_per_launch: int = 1, params: _abi.FrameParams | None = None) -> None:
        p = params if params is not None else self.params()
        self._schedule(p)
        call("rtc8x_trace", C.c_void_p(self.stream), C.byref(p), _abi.ENTRY_PATH_TRACE, int(samples_per_launch))
        self.samples += int(samples_per_launch)

    def image(self) -> np.ndarray:
        cfg = self.scene.cfg
        wait_stream(self.stream)
        a = self.accum.download(np.float64, 3 * cfg.width * cfg.height).reshape(cfg.height, cfg.width, 3)
        return a / max(self.samples, 1)

    def counters(self) -> dict:
        wait_stream(self.stream)
        s = self.stats.download(np.uint64, _abi.STAT_COUNT)
        return {"paths": int(s[_abi.STAT_PATHS]), "vertices": int(s[_abi.STAT_VERTICES]),
                "lookups": int(s[_abi.STAT_LOOKUPS]), "scatters": int(s[_abi.STAT_SCATTER])}

    def close(self) -> None:
        for b in (self.rng, self.accum, self.vertices, self.stats, self.cost, self.order, self.order_scratch):
            if b is not None:
                b.free()
