"""Several instances in one persistent launch (rtc8x_trace_batch) and the autotuner of their scheduling.

An instance is a PathTracer (plain entry) or an NrcRenderer (the NRC render entry), each with its own
scene, image and buffers.  `autotune` times every instance's own launch under each candidate setting of
the tracer's scheduling knobs (the event threshold: how many eighths of a wavefront's active lanes park
before their vertex events run; and the shape of the 64-pixel tiles lanes take), keeps each instance's
fastest setting, and returns the fastest times as the workgroup shares of the batched launch, so that
the instances finish together; `pick_batch_setting` then checks those choices against the plain ones on
the batched launch.  The knobs change the schedule, never the results."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _abi
from ._lib import call
from .device import EventTimer, synchronize
from .nrc import NrcRenderer, round_up
from .tracer import PathTracer


def _launch(params: Sequence[_abi.FrameParams], samples: Sequence[int], entry: int, shares, stream) -> None:
    n = len(params)
    arr = (_abi.FrameParams * n)(*params)
    ns = (C.c_uint32 * n)(*[int(k) for k in samples])
    sh = None if shares is None else (C.c_float * n)(*[float(x) for x in shares])
    call("rtc8x_trace_batch", C.c_void_p(stream), arr, ns, sh, n, entry)


def trace_plain(tracers: Sequence[PathTracer], samples: Sequence[int], shares=None, stream=None) -> None:
    """PathTracer.trace(samples[i]) for every tracer, in one launch."""
    params = []
    for t in tracers:
        p = t.params()
        t._schedule(p)
        params.append(p)
    _launch(params, samples, _abi.ENTRY_PATH_TRACE, shares, stream)
    for t, k in zip(tracers, samples):
        t.samples += int(k)


def render_nrc(renderers: Sequence[NrcRenderer], samples: Sequence[int], shares=None, stream=None) -> None:
    """NrcRenderer.render_samples(samples[i]) for every renderer, with the traces in one launch (the
    inference and accumulation stay per renderer: every instance has its own cache)."""
    params = []
    for r, k in zip(renderers, samples):
        assert 1 <= k <= r.samples_per_launch
        p = r.params()
        p.nrc.num_samples = int(k)
        if r.have_cost:
            call("rtc8x_order_by_tile_cost", C.c_void_p(stream), C.c_void_p(r.cost.ptr), r.width, r.height,
                 C.c_void_p(r.order.ptr), C.c_void_p(r.order_scratch.ptr))
            p.work_order = r.order.ptr
        p.pixel_cost = r.cost.ptr
        r.have_cost = True
        params.append(p)
    _launch(params, samples, _abi.ENTRY_PATH_TRACE_WITH_NRC, shares, stream)
    for r, p, k in zip(renderers, params, samples):
        nq = round_up(r.width * r.height, 128) * int(k)
        r.cache.infer_render(stream, p)
        call("rtc8x_nrc_accumulate", C.c_void_p(stream), C.byref(p))
        r.samples += int(k)
        r.stats.inference_queries += nq


# (event_eighths, tile_shape) candidates: events from 2/8 to 5/8 of the active lanes; 8x8, 16x4, 4x16 tiles
KNOB_CANDIDATES = [(ev, shape) for ev in (2, 3, 4, 5) for shape in (0, 1, 3)]
TILE_SHAPES = {0: "8x8", 1: "16x4", 2: "32x2", 3: "4x16"}


def _time(fn, repeats: int, stream) -> float:
    fn()
    best = float("inf")
    for _ in range(repeats):
        t = EventTimer()
        t.start(stream)
        fn()
        t.stop(stream)
        synchronize()
        best = min(best, t.elapsed_ms())
        t.destroy()
    return best


def autotune(launch_one, count: int, candidates=KNOB_CANDIDATES, repeats: int = 2, stream=None):
    """launch_one(i, event_eighths, tile_shape) launches instance i alone with those knobs.  Returns
    (shares, knobs): per instance the time of its fastest candidate and that candidate."""
    shares, knobs = [], []
    for i in range(count):
        timed = [(_time(lambda: launch_one(i, ev, shape), repeats, stream), (ev, shape)) for ev, shape in candidates]
        best = min(timed)
        shares.append(best[0])
        knobs.append(best[1])
    return shares, knobs


def apply_knobs(params: _abi.FrameParams, knobs) -> _abi.FrameParams:
    params.event_eighths, params.tile_shape = int(knobs[0]), int(knobs[1])
    return params


def pick_batch_setting(launch_batch, settings, repeats: int = 3, stream=None):
    """The autotuner's last step, on the batched launch itself: times launch_batch(shares, knobs) for
    every candidate (shares, knobs) in `settings` (the best of `repeats`, after one untimed launch) and
    returns (shares, knobs, ms) of the fastest.  Solo launch times make shares only approximately: in the
    batch the instances share the device and wavefronts move on to unfinished instances, so the solo
    choices are checked against the plain ones (shares by pixels, default knobs) there."""
    best = None
    for shares, knobs in settings:
        ms = _time(lambda: launch_batch(shares, knobs), repeats, stream)
        if best is None or ms < best[2]:
            best = (shares, knobs, ms)
    return best


def autotune_shares(launch_one, count: int, repeats: int = 2, stream=None) -> list[float]:
    """Times launch_one(i) for every instance i (the best of `repeats`, after one untimed launch) and
    returns the times: the shares under which the instances of one batched launch finish together."""
    return [_time(lambda: launch_one(i), repeats, stream) for i in range(count)]
