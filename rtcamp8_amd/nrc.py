"""Neural radiance cache: the host-side mirror of the reference's NeuralRadianceCache class
(renderer/network_interface.h:17-31) and of the NRC branch of its frame loop
(renderer/renderer_main.cpp:1554-1638), driving the gfx950 library through the C ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._lib import call, lib
from .device import DeviceBuffer, Event, Stream, wait_stream
from .scene import DeviceScene

TRAIN_BATCH = _abi.NUM_TRAINING_DATA_PER_FRAME // 4  # renderer_main.cpp:1595: 4 steps of 16384 per iteration


class Mt19937:
    """std::mt19937 (32-bit Mersenne Twister), used by the reference as perFrameRng (renderer_main.cpp:1500)."""

    def __init__(self, seed: int):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def __call__(self) -> int:
        if self.idx >= 624:
            for i in range(624):
                y = (self.mt[i] & 0x80000000) | (self.mt[(i + 1) % 624] & 0x7FFFFFFF)
                v = self.mt[(i + 397) % 624] ^ (y >> 1)
                if y & 1:
                    v ^= 0x9908B0DF
                self.mt[i] = v
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


class NeuralRadianceCache:
    """initialize / finalize / infer / train, as network_interface.h declares them."""

    def __init__(self):
        self.handle = None

    def initialize(self, num_hidden_layers: int = 2, learning_rate: float = 1e-2, precision: int = _abi.NRC_FP16,
                   seed: int = 0x9E3779B97F4A7C15, ema_decay: float = 0.99,
                   encoding: int = _abi.NRC_HASH_GRID, loss_scale: float = 0.0) -> None:
        """initialize(posEnc, numHiddenLayers, learningRate) (network_interface.cu:50-134); Adam epsilon
        1e-15 with the hash grid and 1e-8 with triangle waves, as the reference configures them.
        loss_scale: the initial loss scale (0: 128, tiny-cuda-nn's; dynamic, with FP16's range as the
        overflow rule in every precision; a large one skips a fresh network's outlying first steps)."""
        eps = 1e-15 if encoding == _abi.NRC_HASH_GRID else 1e-8
        d = _abi.NrcDesc(num_hidden_layers, learning_rate, 0.9, 0.99, eps, 1e-6, ema_decay, precision, seed, encoding,
                         float(loss_scale))
        h = C.c_void_p()
        call("rtc8x_nrc_create", C.byref(d), C.byref(h))
        self.handle = h.value
        self.num_hidden_layers = num_hidden_layers
        self.encoding = encoding

    def finalize(self) -> None:
        if self.handle:
            call("rtc8x_nrc_destroy", C.c_void_p(self.handle))
            self.handle = None

    def infer(self, stream, input_ptr: int, num: int, output_ptr: int) -> None:
        call("rtc8x_nrc_infer", C.c_void_p(self.handle), C.c_void_p(stream), C.c_void_p(input_ptr), int(num),
             C.c_void_p(output_ptr))

    def infer_render(self, stream, params: _abi.FrameParams) -> None:
        """The render inference of a PATH_TRACE_WITH_NRC launch's samples, at the slots with a query."""
        call("rtc8x_nrc_infer_render", C.c_void_p(self.handle), C.c_void_p(stream), C.byref(params))

    def infer_training(self, stream, params: _abi.FrameParams) -> None:
        """Inference of a training iteration's suffix queries, sized on the device (no host wait)."""
        call("rtc8x_nrc_infer_training", C.c_void_p(self.handle), C.c_void_p(stream), C.byref(params))

    def train(self, stream, input_ptr: int, target_ptr: int, num: int, want_loss: bool = False):
        loss = C.c_float(0.0)
        call("rtc8x_nrc_train", C.c_void_p(self.handle), C.c_void_p(stream), C.c_void_p(input_ptr),
             C.c_void_p(target_ptr), int(num), C.byref(loss) if want_loss else None)
        return float(loss.value) if want_loss else None

    def loss_scale(self) -> float:
        v = C.c_float(0.0)
        call("rtc8x_nrc_get_loss_scale", C.c_void_p(self.handle), C.byref(v))
        return float(v.value)

    def num_params(self) -> int:
        return int(lib().rtc8x_nrc_num_params(C.c_void_p(self.handle)))

    def get_weights(self, which: int = 0) -> np.ndarray:
        w = np.empty(self.num_params(), dtype=np.float32)
        call("rtc8x_nrc_get_weights", C.c_void_p(self.handle), which, w.ctypes.data_as(C.c_void_p), w.size)
        return w

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        call("rtc8x_nrc_set_weights", C.c_void_p(self.handle), w.ctypes.data_as(C.c_void_p), w.size)

    def gradient(self, stream, input_ptr: int, target_ptr: int, num: int) -> np.ndarray:
        g = DeviceBuffer(4 * self.num_params())
        call("rtc8x_nrc_gradient", C.c_void_p(self.handle), C.c_void_p(stream), C.c_void_p(input_ptr),
             C.c_void_p(target_ptr), int(num), C.c_void_p(g.ptr))
        wait_stream(stream)
        out = g.download(np.float32, self.num_params())
        g.free()
        return out


def round_up(x: int, m: int) -> int:
    return (x + m - 1) // m * m


class NrcBuffers:
    """Screen-size and training buffers of the NRC path (renderer_main.cpp:80-129, 197-230)."""

    def __init__(self, width: int, height: int, shuffler_seed: int = 471313181, samples: int = 1):
        """`samples`: image slices of the render buffers (samples one PATH_TRACE_WITH_NRC launch traces)."""
        n = width * height
        nb_inf = round_up(n, 128)
        self.width, self.height, self.samples = width, height, samples
        self.num_training_data = DeviceBuffer.zeros(8)
        self.train_image_size = DeviceBuffer.zeros(16)
        self.target_min_max = DeviceBuffer.zeros(4 * 12)
        self.target_avg = DeviceBuffer.zeros(4 * 6)
        self.inference_queries = DeviceBuffer.zeros(56 * nb_inf * samples)
        self.inference_terminals = DeviceBuffer.zeros(16 * n * samples)
        self.train_suffix_terminals = DeviceBuffer.zeros(16 * n)
        self.inferred_radiance = DeviceBuffer.zeros(12 * nb_inf * samples)
        self.per_frame_contribution = DeviceBuffer.zeros(12 * n * samples)
        self.train_queries = [DeviceBuffer.zeros(56 * _abi.TRAIN_BUFFER_SIZE) for _ in range(2)]
        self.train_targets = [DeviceBuffer.zeros(12 * _abi.TRAIN_BUFFER_SIZE) for _ in range(2)]
        self.train_vertex_info = DeviceBuffer.zeros(24 * _abi.TRAIN_BUFFER_SIZE)
        # LCG states: the i-th slot holds the state after i + 1 steps from the seed (renderer_main.cpp:220-229)
        s = shuffler_seed
        states = np.empty(_abi.NUM_TRAINING_DATA_PER_FRAME, dtype=np.uint32)
        for i in range(states.size):
            s = ((s * 1103515245 + 12345) & 0xFFFFFFFF) % (1 << 31)
            states[i] = s
        self.data_shufflers = DeviceBuffer.from_array(states)
        self.scratch = DeviceBuffer.zeros(lib().rtc8x_nrc_scratch_bytes(width, height))

    def fill(self, nrc: _abi.NrcBuffers) -> None:
        nrc.max_num_training_suffixes = self.width * self.height
        nrc.num_training_data = self.num_training_data.ptr
        nrc.train_image_size = self.train_image_size.ptr
        nrc.target_min_max = self.target_min_max.ptr
        nrc.target_avg = self.target_avg.ptr
        nrc.inference_queries = self.inference_queries.ptr
        nrc.inference_terminals = self.inference_terminals.ptr
        nrc.train_suffix_terminals = self.train_suffix_terminals.ptr
        nrc.inferred_radiance = self.inferred_radiance.ptr
        nrc.per_frame_contribution = self.per_frame_contribution.ptr
        nrc.train_queries[0], nrc.train_queries[1] = self.train_queries[0].ptr, self.train_queries[1].ptr
        nrc.train_targets[0], nrc.train_targets[1] = self.train_targets[0].ptr, self.train_targets[1].ptr
        nrc.train_vertex_info = self.train_vertex_info.ptr
        nrc.data_shufflers = self.data_shufflers.ptr
        nrc.scratch = self.scratch.ptr

    def free(self) -> None:
        for v in list(self.__dict__.values()):
            if isinstance(v, DeviceBuffer):
                v.free()
            elif isinstance(v, list):
                for b in v:
                    if isinstance(b, DeviceBuffer):
                        b.free()


@dataclass
class NrcStats:
    train_iterations: int = 0
    train_steps: int = 0
    inference_queries: int = 0
    last_loss: float | None = None


class NrcRenderer:
    """Frame loop of the NRC-accelerated tracer (renderer_main.cpp:1554-1638): per frame, training
    iterations (8 on the first frame, then 1) of prepare -> trace training paths -> infer terminal
    queries -> propagate -> shuffle -> 4 x train(16384); then per sample trace with NRC -> infer ->
    accumulate.  The plain tracer (rtcamp8_amd.tracer) stays the unbiased reference.

    The traces run on a second stream of the renderer's own.  A trace that directly follows a training
    iteration starts after that iteration's shuffle and runs beside its four network steps: the trace
    neither reads the network nor touches the shuffled training buffers the steps read, and the
    network is read again only by the inference that waits for the trace.  Every other trace waits for
    all work queued on the renderer's stream.  Results equal those of one stream."""

    def __init__(self, scene: DeviceScene, num_hidden_layers: int = 2, learning_rate: float = 1e-2,
                 precision: int = _abi.NRC_FP16, stream=None, radiance_scale: float = 0.0, frame_rng_seed: int = 72139121,
                 encoding: int = _abi.NRC_HASH_GRID, ema_decay: float = 0.99, first_frame_iterations: int = 8,
                 samples_per_launch: int = 4, train_tracking: int | None = None, loss_scale: float = 0.0,
                 overlap: bool = True):
        """samples_per_launch: samples traced into the cache by one launch (and inferred and accumulated
        together); the sums equal those of one-sample launches.  train_tracking: the tracker of the
        training-data launches (None: the scene's); both trackers give training data of the same
        distribution, and the global one runs the training entry faster (its local variant holds more
        registers and runs two waves per SIMD).  loss_scale: the cache's initial loss scale (0: 128).
        overlap: traces that follow a training iteration run beside its network steps (False: every
        trace waits for all earlier work)."""
        cfg = scene.cfg
        self.scene = scene
        self.stream = stream
        self.width, self.height = cfg.width, cfg.height
        n = cfg.width * cfg.height
        self.rng = DeviceBuffer(8 * n)
        self.accum = DeviceBuffer(24 * n)
        self.stats_buf = DeviceBuffer.zeros(8 * _abi.STAT_COUNT)
        self.samples_per_launch = max(1, int(samples_per_launch))
        self.buffers = NrcBuffers(cfg.width, cfg.height, samples=self.samples_per_launch)
        # cost-ordered scheduling of the render launches (the full-image entry)
        self.cost = DeviceBuffer.zeros(4 * n)
        self.order = DeviceBuffer(4 * n)
        self.order_scratch = DeviceBuffer(lib().rtc8x_order_tiles_scratch_bytes(cfg.width, cfg.height))
        self.have_cost = False
        self.cache = NeuralRadianceCache()
        self.cache.initialize(num_hidden_layers, learning_rate, precision, encoding=encoding, ema_decay=ema_decay,
                              loss_scale=loss_scale)
        self.first_frame_iterations = first_frame_iterations
        self.frame_rng = Mt19937(frame_rng_seed)
        self.buf_idx = 1  # the first iteration toggles to 0 (renderer_main.cpp:1558-1559)
        self.radiance_scale = radiance_scale
        self.stats = NrcStats()
        self.samples = 0
        self.frame = 0
        self.knobs = (0, 0)  # (event_eighths, tile_shape) of the trace kernel's schedule; 0 = defaults
        self.train_tracking = train_tracking
        self.side = Stream()
        self.gate = Event()    # recorded after a training iteration's shuffle, before its network steps
        self.fence = Event()
        self.traced = Event()
        self.gate_open = False  # no launch on the renderer's buffers since the gate was recorded
        self.overlap = overlap
        call("rtc8x_seed_rng", C.c_void_p(self.rng.ptr), cfg.width, cfg.height, C.c_uint64(cfg.rng_seed))
        self.accum.fill(0)

    def reset(self, seed: int) -> None:
        """Restarts the accumulation from new per-pixel seeds (the network keeps its training)."""
        wait_stream(self.stream)
        self.gate_open = False
        cfg = self.scene.cfg
        call("rtc8x_seed_rng", C.c_void_p(self.rng.ptr), cfg.width, cfg.height, C.c_uint64(seed))
        self.accum.fill(0)
        self.samples = 0

    def params(self) -> _abi.FrameParams:
        """The frame parameters of a launch on the renderer's buffers; closes the gate, so that the next
        trace waits for everything queued before it."""
        self.gate_open = False
        p = self.scene.params()
        p.rng_states = self.rng.ptr
        p.accum = self.accum.ptr
        p.stats = self.stats_buf.ptr
        self.buffers.fill(p.nrc)
        p.nrc.nrc_buffer_index = self.buf_idx
        p.nrc.radiance_scale = self.radiance_scale
        p.nrc.path_termination_factor = 0.01
        p.event_eighths, p.tile_shape = self.knobs
        return p

    def _trace(self, after_gate: bool, issue) -> None:
        """Queues issue(side stream) after the gate (after_gate) or after all work on the renderer's
        stream; the renderer's stream then waits for it."""
        side = self.side.handle
        if after_gate:
            self.gate.make_wait(side)
        else:
            self.fence.record(self.stream)
            self.fence.make_wait(side)
        issue(C.c_void_p(side))
        self.traced.record(side)
        self.traced.make_wait(self.stream)

    def train_iteration(self, is_new_sequence: bool, want_loss: bool = False):
        s = C.c_void_p(self.stream)
        after_gate = self.gate_open and self.overlap
        self.buf_idx = (self.buf_idx + 1) % 2
        p = self.params()
        if self.train_tracking is not None:
            p.tracking = self.train_tracking
        p.nrc.offset_to_select_unbiased_path = self.frame_rng()

        def trace(side):
            call("rtc8x_nrc_prepare", side, C.byref(p), 1 if is_new_sequence else 0)
            call("rtc8x_trace", side, C.byref(p), _abi.ENTRY_GENERATE_TRAINING_DATA, 1)

        self._trace(after_gate, trace)
        # the training size stays on the device: no host synchronisation in the iteration
        self.cache.infer_training(self.stream, p)
        call("rtc8x_nrc_propagate", s, C.byref(p))
        call("rtc8x_nrc_shuffle", s, C.byref(p))
        self.gate.record(self.stream)
        loss = None
        for step in range(4):
            off = step * TRAIN_BATCH
            loss = self.cache.train(self.stream, self.buffers.train_queries[1].ptr + 56 * off,
                                    self.buffers.train_targets[1].ptr + 12 * off, TRAIN_BATCH,
                                    want_loss=want_loss and step == 3)
        self.stats.train_iterations += 1
        self.stats.train_steps += 4
        self.stats.last_loss = loss
        self.gate_open = True
        return loss

    def render_samples(self, k: int = 1) -> None:
        """k samples (k <= samples_per_launch) traced into the cache by one launch, inferred by one
        network call and accumulated in sample order."""
        assert 1 <= k <= self.samples_per_launch
        s = C.c_void_p(self.stream)
        after_gate = self.gate_open and self.overlap
        p = self.params()
        p.nrc.num_samples = k
        ordered = self.have_cost
        if ordered:
            p.work_order = self.order.ptr
        p.pixel_cost = self.cost.ptr
        self.have_cost = True

        def trace(side):
            if ordered:
                call("rtc8x_order_by_tile_cost", side, C.c_void_p(self.cost.ptr), self.width, self.height,
                     C.c_void_p(self.order.ptr), C.c_void_p(self.order_scratch.ptr))
            call("rtc8x_trace", side, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, k)

        self._trace(after_gate, trace)
        nq = round_up(self.width * self.height, 128) * k
        self.cache.infer_render(self.stream, p)
        call("rtc8x_nrc_accumulate", s, C.byref(p))
        self.stats.inference_queries += nq
        self.samples += k

    def render_sample(self) -> None:
        self.render_samples(1)

    def render_frame(self, spp: int, train_iterations: int | None = None) -> None:
        iters = train_iterations if train_iterations is not None else (self.first_frame_iterations if self.frame == 0 else 1)
        for _ in range(iters):
            self.train_iteration(self.frame == 0)
        done = 0
        while done < spp:
            k = min(self.samples_per_launch, spp - done)
            self.render_samples(k)
            done += k
        self.frame += 1

    def image(self) -> np.ndarray:
        wait_stream(self.stream)
        a = self.accum.download(np.float64, 3 * self.width * self.height).reshape(self.height, self.width, 3)
        return a / max(self.samples, 1)

    def counters(self) -> dict:
        wait_stream(self.stream)
        s = self.stats_buf.download(np.uint64, _abi.STAT_COUNT)
        return {"paths": int(s[_abi.STAT_PATHS]), "vertices": int(s[_abi.STAT_VERTICES]),
                "lookups": int(s[_abi.STAT_LOOKUPS])}

    def close(self) -> None:
        wait_stream(self.stream)
        self.side.synchronize()
        for e in (self.gate, self.fence, self.traced):
            e.destroy()
        self.side.destroy()
        self.cache.finalize()
        self.buffers.free()
        for b in (self.rng, self.accum, self.stats_buf, self.cost, self.order, self.order_scratch):
            b.free()
