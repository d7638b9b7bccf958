"""GPU parity of the neural radiance cache: the MFMA network against the FP64 oracle (FP16 / BF16
storage emulated there, so the remaining difference is FP32 accumulation order; FP32 weights against
the unrounded oracle, so the difference is FP32 rounding), the optimiser step
bit for bit, training convergence, and the NRC data path and trace entry points bit for bit."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import HostNrc, download_env, download_grid, oracle, oracle_seed, ptr
from test_oracle_nrc import random_queries, random_weights

pytestmark = pytest.mark.gpu


def _cache(L, precision, encoding):
    from rtcamp8_amd.nrc import NeuralRadianceCache

    c = NeuralRadianceCache()
    c.initialize(num_hidden_layers=L, learning_rate=1e-2, precision=precision, encoding=encoding)
    return c


ENCODINGS = [0, 1]  # triangle waves, hash grid


def test_triangle_encoding_matches_oracle():
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    n = 3000
    q = random_queries(n, seed=9)
    dq, de = DeviceBuffer.from_array(q), DeviceBuffer(4 * 64 * n)
    call("rtc8x_nrc_encode", None, C.c_void_p(dq.ptr), n, C.c_void_p(de.ptr))
    got = de.download(np.float32, 64 * n).astype(np.float64)
    ref = np.zeros(64 * n)
    oracle().oracle_nrc_encode(ptr(q), n, ptr(ref))
    assert np.max(np.abs(got - ref)) < 2e-6


def test_hash_encoding_matches_oracle():
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    L, n = 2, 3000
    q = random_queries(n, seed=9)
    q[:5, 0:3] = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [0.5, 0.5, 0.5], [1e-7, 1 - 1e-7, 0.25]]  # corners, faces
    w = random_weights(L, seed=19, encoding=1)
    c = _cache(L, 0, 1)
    c.set_weights(w)
    dq, de = DeviceBuffer.from_array(q), DeviceBuffer(4 * 64 * n)
    call("rtc8x_nrc_encode_net", C.c_void_p(c.handle), None, C.c_void_p(dq.ptr), n, C.c_void_p(de.ptr), 0)
    got = de.download(np.float32, 64 * n).astype(np.float64).reshape(n, 64)
    ref = np.zeros(64 * n)
    oracle().oracle_nrc_encode_ex(ptr(w), L, 1, ptr(q), n, ptr(ref))
    ref = ref.reshape(n, 64)
    # FP32 interpolation on the GPU: a cell's weights carry ~1e-7 relative rounding per level
    assert np.max(np.abs(got - ref)) < 2e-6
    c.finalize()


@pytest.mark.parametrize("encoding", ENCODINGS)
@pytest.mark.parametrize("L,precision,mode,tol", [(2, 0, 1, 4e-3), (5, 0, 1, 4e-3), (2, 1, 2, 2e-2), (5, 1, 2, 2e-2),
                                                   (2, 3, 0, 1e-4), (5, 3, 0, 1e-4)])
def test_inference_matches_oracle(L, precision, mode, tol, encoding):
    from rtcamp8_amd.device import DeviceBuffer

    n = 1000  # not a multiple of the 256-query workgroup: exercises the tail
    q = random_queries(n, seed=10)
    w = random_weights(L, seed=11, scale=1.2, encoding=encoding, table_scale=1.0)
    c = _cache(L, precision, encoding)
    c.set_weights(w)
    dq, do = DeviceBuffer.from_array(q), DeviceBuffer(12 * n)
    c.infer(None, dq.ptr, n, do.ptr)
    got = do.download(np.float32, 3 * n).astype(np.float64)
    ref = np.zeros(3 * n)
    oracle().oracle_nrc_forward_ex(ptr(w), L, encoding, ptr(q), n, ptr(ref), mode)
    rms = np.sqrt(np.mean(ref ** 2))
    assert rms > 1e-3
    assert np.max(np.abs(got - ref)) < tol * rms, np.max(np.abs(got - ref)) / rms
    c.finalize()


def test_loss_scale_keeps_small_fp16_gradients():
    """The loss scale against FP16's small-number range.  With the output layer's weights scaled by 1e-2
    the outputs p are ~1e-2, and targets t = p + d put the residual at d = 3n x 1e-7 (n = 512) or d x 2^12.
    The small residual's per-sample gradients of the mean loss (~2e-5) and the hidden layers' input
    gradients behind them (~2e-7) are FP16 subnormals with few significant bits; 2^12 times larger they are
    normal FP16 numbers.
    The gradient is linear in the residual, so g(d) x 2^12 should equal g(d x 2^12).  Unscaled (loss scale
    1) the hidden layers' weight gradients for the small residual lose most of their bits; with the
    default scale (128) they are computed in FP16's normal range.  Measured: relative error 0.856 unscaled,
    0.0064 with the scale."""
    from rtcamp8_amd.device import DeviceBuffer

    L, n, k = 3, 512, 2.0 ** 12
    q = random_queries(n, seed=21)
    w = random_weights(L, seed=22, scale=1.2, encoding=0, table_scale=1.0)
    w[L * 4096:L * 4096 + 16 * 64] *= np.float32(1e-2)
    dq = DeviceBuffer.from_array(q)
    sign = np.where(np.random.default_rng(23).random((n, 3)) < 0.5, -1.0, 1.0).astype(np.float32)
    res = {}
    for scale in (1.0, 0.0):
        c = _cache_scaled(L, 0, 0, scale)
        c.set_weights(w)
        dp = DeviceBuffer(12 * n)
        c.infer(None, dq.ptr, n, dp.ptr)
        p = dp.download(np.float32, 3 * n).reshape(n, 3)
        grads = []
        for d in (3 * n * 1e-7, 3 * n * 1e-7 * k):
            t = (p + np.float32(d) * sign).astype(np.float32)
            dt = DeviceBuffer.from_array(t)
            grads.append(c.gradient(None, dq.ptr, dt.ptr, n).astype(np.float64))
            dt.free()
        g_small, g_big = grads
        hidden = slice(0, L * 4096)
        res[scale] = float(np.linalg.norm(g_small[hidden] * k - g_big[hidden]) / np.linalg.norm(g_big[hidden]))
        print(f"[loss scale] initial {c.loss_scale():g}: relative error of the hidden layers' small-residual "
              f"gradient {res[scale]:.4f} (output |p| mean {np.abs(p).mean():.3g})", flush=True)
        dp.free()
        c.finalize()
    dq.free()
    assert res[0.0] < LOSS_SCALE_ERR_SCALED, res
    assert res[1.0] > 4.0 * res[0.0], res


LOSS_SCALE_ERR_SCALED = 0.011


def _cache_scaled(L, precision, encoding, loss_scale):
    from rtcamp8_amd.nrc import NeuralRadianceCache

    c = NeuralRadianceCache()
    c.initialize(num_hidden_layers=L, learning_rate=1e-2, precision=precision, encoding=encoding, loss_scale=loss_scale)
    return c


def test_loss_scale_backs_off_after_overflow():
    """The dynamic loss scale: started at 2^24 with targets far from the outputs (per-sample gradients of
    the mean loss up to ~0.2), the scaled FP16 loss gradient overflows; those steps are skipped and the
    scale halves until it fits (at most 2^20), after which training goes on and the loss falls.  Measured:
    scale 2^18 after 31 steps, loss 871 -> 0.53."""
    from rtcamp8_amd.device import DeviceBuffer

    n = 4096
    q = random_queries(n, seed=31)
    t = (10.0 * np.random.default_rng(32).random((n, 3))).astype(np.float32)
    dq, dt = DeviceBuffer.from_array(q), DeviceBuffer.from_array(t)
    c = _cache_scaled(2, 0, 0, 2.0 ** 24)
    first = c.train(None, dq.ptr, dt.ptr, n, want_loss=True)
    w0 = c.get_weights()
    for _ in range(30):
        last = c.train(None, dq.ptr, dt.ptr, n, want_loss=True)
    s = c.loss_scale()
    print(f"[loss scale] after 31 steps: scale {s:g}, loss {first:.4g} -> {last:.4g}", flush=True)
    assert 1.0 <= s <= 2.0 ** 20, s
    assert np.isfinite(c.get_weights()).all()
    assert np.any(c.get_weights() != w0)
    assert last < first, (first, last)
    c.finalize()
    dq.free()
    dt.free()


FP8_REL_RMS = {2: 0.085, 5: 0.155}
FP8_MAX_OF_RMS = {2: 0.33, 5: 0.85}


@pytest.mark.parametrize("encoding", ENCODINGS)
@pytest.mark.parametrize("L", [2, 5])
def test_fp8_inference_matches_oracle(L, encoding):
    """E4M3 inference (BF16 training) against the exact FP64 forward of the same weights.  E4M3 keeps 3
    mantissa bits: each operand carries up to 2^-4 relative rounding (RMS about 3.6 %), and a 64-term
    dot product of random-sign terms keeps about 5 % of it per layer.  Measured with activations kept in
    E4M3 between layers (per-tile scales from the FP32 accumulators): relative RMS error 0.064 / 0.058 at
    L = 2 and 0.094 / 0.119 at L = 5 (triangle waves / hash grid), max 0.25 / 0.20 and 0.42 / 0.65 of the
    RMS.  Tolerance: about 1.3x the larger measured value at each depth."""
    from rtcamp8_amd.device import DeviceBuffer

    n = 1000
    q = random_queries(n, seed=10)
    w = random_weights(L, seed=11, scale=1.2, encoding=encoding, table_scale=1.0)
    c = _cache(L, 2, encoding)
    c.set_weights(w)
    dq, do = DeviceBuffer.from_array(q), DeviceBuffer(12 * n)
    c.infer(None, dq.ptr, n, do.ptr)
    got = do.download(np.float32, 3 * n).astype(np.float64)
    ref = np.zeros(3 * n)
    oracle().oracle_nrc_forward_ex(ptr(w), L, encoding, ptr(q), n, ptr(ref), 0)
    rms = np.sqrt(np.mean(ref ** 2))
    err = np.sqrt(np.mean((got - ref) ** 2)) / rms
    print(f"[fp8] L={L} encoding={encoding}: relative RMS error {err:.4f}, max {np.max(np.abs(got - ref)) / rms:.4f}")
    assert rms > 1e-3 and np.all(np.isfinite(got))
    assert err < FP8_REL_RMS[L], err
    assert np.max(np.abs(got - ref)) < FP8_MAX_OF_RMS[L] * rms
    c.finalize()


@pytest.mark.parametrize("encoding", ENCODINGS)
@pytest.mark.parametrize("L,precision,mode,tol", [(2, 0, 1, 5e-3), (5, 0, 1, 5e-3), (3, 1, 2, 3e-2), (3, 3, 0, 1e-4)])
def test_gradient_matches_oracle(L, precision, mode, tol, encoding):
    from rtcamp8_amd.device import DeviceBuffer

    n = 700
    q = random_queries(n, seed=12)
    t = np.random.default_rng(13).random((n, 3)).astype(np.float32)
    w = random_weights(L, seed=14, scale=1.2, encoding=encoding, table_scale=1.0)
    P = L * 4096 + 16 * 64
    c = _cache(L, precision, encoding)
    c.set_weights(w)
    dq, dt = DeviceBuffer.from_array(q), DeviceBuffer.from_array(t)
    got = c.gradient(None, dq.ptr, dt.ptr, n).astype(np.float64)
    ref = np.zeros(w.size)
    oracle().oracle_nrc_gradient_ex(ptr(w), L, encoding, ptr(q), ptr(t), n, ptr(ref), mode)
    ref /= 3 * n  # the oracle sums over the samples; the library returns the mean loss's gradient
    sections = [slice(layer * 4096, (layer + 1) * 4096 if layer < L else layer * 4096 + 3 * 64) for layer in range(L + 1)]
    if encoding == 1:
        sections.append(slice(P, w.size))  # the hash tables
    for k, sl in enumerate(sections):  # every layer on its own, so a small layer cannot hide in the total
        err = np.linalg.norm(got[sl] - ref[sl]) / np.linalg.norm(ref[sl])
        assert err < tol, (k, err)
    exact = np.zeros(w.size)
    oracle().oracle_nrc_gradient_ex(ptr(w), L, encoding, ptr(q), ptr(t), n, ptr(exact), 0)
    exact /= 3 * n
    assert np.linalg.norm(got - exact) / np.linalg.norm(exact) < 10 * tol
    assert np.all(got[L * 4096 + 3 * 64:P] == 0)
    c.finalize()


@pytest.mark.parametrize("encoding", ENCODINGS)
def test_adam_step_bit_exact(encoding):
    from rtcamp8_amd.device import DeviceBuffer

    L, n = 2, 2048
    q = random_queries(n, seed=15)
    t = np.random.default_rng(16).random((n, 3)).astype(np.float32)
    w = random_weights(L, seed=17, scale=1.0, encoding=encoding)
    c = _cache(L, 0, encoding)
    c.set_weights(w)
    eps = 1e-15 if encoding == 1 else 1e-8
    dq, dt = DeviceBuffer.from_array(q), DeviceBuffer.from_array(t)
    g = c.gradient(None, dq.ptr, dt.ptr, n)
    c.train(None, dq.ptr, dt.ptr, n)
    master = c.get_weights(0)
    ema = c.get_weights(1)
    om, m, v, oe = w.copy(), np.zeros_like(w), np.zeros_like(w), np.zeros_like(w)
    oracle().oracle_nrc_adam(ptr(om), ptr(m), ptr(v), ptr(oe), ptr(g), w.size, 1e-2, 0.9, 0.99, eps, 1e-6, 0.99, 1)
    # the debiasing divisor is 1 - decay^t in double from the FP32 decay, rounded to FP32
    debias = np.float32(1.0 - float(np.float32(0.99)))
    P = L * 4096 + 16 * 64
    assert np.array_equal(master[:P].view(np.uint32), om[:P].view(np.uint32))
    assert np.array_equal(ema[:P].view(np.uint32), (oe[:P] / debias).view(np.uint32))
    if encoding == 1:
        # the tables' gradient is scattered with atomics, so the step's own gradient can differ from `g`
        # in the last bits; Adam's first step moves each entry by about lr * sign(g)
        diff = np.abs(master[P:] - om[P:]) > 1e-6 * np.maximum(np.abs(om[P:]), 1e-3)
        assert diff.mean() < 1e-3, diff.mean()
    # the working copies written by the optimiser step: the gradient at the new weights matches the
    # oracle's at the new master weights
    g2 = c.gradient(None, dq.ptr, dt.ptr, n).astype(np.float64)
    ref = np.zeros(w.size)
    oracle().oracle_nrc_gradient_ex(ptr(master), L, encoding, ptr(q), ptr(t), n, ptr(ref), 1)
    ref /= 3 * n
    assert np.linalg.norm(g2 - ref) / np.linalg.norm(ref) < 5e-3
    c.finalize()


@pytest.mark.parametrize("encoding", ENCODINGS)
@pytest.mark.parametrize("precision", [0, 1, 3])
def test_training_reduces_loss(precision, encoding):
    """Online training on a smooth synthetic radiance field drops the relative L2 loss by > 10x.  The
    field is mirror-symmetric in each position coordinate so that it is representable by both
    encodings."""
    from rtcamp8_amd.device import DeviceBuffer

    L, n = 3, 16384
    c = _cache(L, precision, encoding)
    losses = []
    for step in range(300):
        q = random_queries(n, seed=100 + step)
        x = q[:, 0:3].astype(np.float64)
        t = np.stack([0.5 + 0.4 * np.cos(2 * np.pi * x[:, 0]),
                      0.4 + 0.3 * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2]),
                      0.3 + 0.2 * np.cos(4 * np.pi * x[:, 2]) + 0.1 * q[:, 3] / np.pi], axis=1).astype(np.float32)
        dq, dt = DeviceBuffer.from_array(q), DeviceBuffer.from_array(t)
        loss = c.train(None, dq.ptr, dt.ptr, n, want_loss=(step % 25 == 0 or step == 299))
        if loss is not None:
            losses.append(loss)
        dq.free()
        dt.free()
    assert np.isfinite(losses).all()
    assert losses[-1] < losses[0] / 10, losses
    c.finalize()


def _nrc_case(width=96, height=64, tracking=0, density=None):
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.nrc import NrcRenderer

    cfg = sc.CONFIGS["dense128"].with_(width=width, height=height, tracking=tracking,
                                       fbm=sc.FbmConfig(resolution=64, seed=0x5EED0001))
    if density is not None:
        cfg = cfg.with_(density_coeff=density)
    ds = sc.DeviceScene(cfg)
    r = NrcRenderer(ds, num_hidden_layers=2)
    return cfg, ds, r


def test_fp8_render_matches_bf16_render():
    """An NRC render with E4M3 inference against the same render with BF16 inference: one cache trained
    online in BF16 on the dense cloud (96 x 64, frame 0's 8 iterations), its inference weights copied into
    a BF16 and a BF16_FP8 cache, each rendering 4 spp from the same seeds without further training.  The
    paths are identical, so the images differ only by the cache's E4M3 rounding at the path ends.
    Measured: relative RMS difference 0.012, mean shift 0.0048; bounds about 1.3x those."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd.nrc import NrcRenderer

    cfg, ds, r = _nrc_case()
    r.render_frame(spp=1)
    w = r.cache.get_weights(1)  # the EMA weights inference uses
    r.close()
    imgs = {}
    for prec in (_abi.NRC_BF16, _abi.NRC_BF16_FP8):
        rr = NrcRenderer(ds, num_hidden_layers=2, precision=prec)
        rr.cache.set_weights(w)
        rr.reset(0xF8F8)
        rr.render_frame(spp=4, train_iterations=0)
        imgs[prec] = rr.image()
        rr.close()
    a, b = imgs[_abi.NRC_BF16], imgs[_abi.NRC_BF16_FP8]
    assert np.isfinite(a).all() and np.isfinite(b).all()
    rel = float(np.sqrt(np.mean((b - a) ** 2)) / np.sqrt(np.mean(a ** 2)))
    mean_shift = float(abs(b.mean() - a.mean()) / a.mean())
    print(f"[fp8 render] relative RMS difference {rel:.5f}, mean shift {mean_shift:.5f}", flush=True)
    assert 0.0 < rel < FP8_RENDER_REL_RMS, rel
    assert mean_shift < FP8_RENDER_MEAN_SHIFT, mean_shift
    ds.close()


def test_frame_loop_overlap_matches_one_stream():
    """The frame loop with each trace beside the preceding training iteration's network steps (the
    default) against the loop with every trace after all earlier work: images, weights and counters equal
    bit for bit over three frames (frame 0's 8 iterations, then 1 per frame, 2 spp each), on the caller's
    stream and on the null stream.  Triangle-wave encoding, whose training is deterministic."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.device import Stream
    from rtcamp8_amd.nrc import NrcRenderer

    cfg = sc.CONFIGS["dense128"].with_(width=96, height=64, fbm=sc.FbmConfig(resolution=64, seed=0x5EED0001))
    ds = sc.DeviceScene(cfg)
    st = Stream()
    for stream in (st.handle, None):
        out = {}
        for overlap in (True, False):
            print(f"[frame loop] overlap {overlap}, stream {'own' if stream else 'null'}", flush=True)
            r = NrcRenderer(ds, num_hidden_layers=2, encoding=_abi.NRC_TRIANGLE_WAVE, stream=stream, overlap=overlap,
                            samples_per_launch=2, radiance_scale=10.0)
            r.reset(0x0F0F)
            for _ in range(3):
                r.render_frame(spp=2)
            out[overlap] = (r.image(), r.cache.get_weights(0), r.cache.get_weights(1), r.counters())
            r.close()
        (ia, wa, ea, ca), (ib, wb, eb, cb) = out[True], out[False]
        assert np.isfinite(ia).all() and ia.mean() > 0
        assert np.array_equal(ia.view(np.uint64), ib.view(np.uint64))
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
        assert np.array_equal(ea.view(np.uint32), eb.view(np.uint32))
        assert ca == cb and ca["paths"] > 0
    st.destroy()
    ds.close()


FP8_RENDER_REL_RMS = 0.016
FP8_RENDER_MEAN_SHIFT = 0.0065


def _host_params(cfg, ds, r, host_nrc, rng, accum):
    from rtcamp8_amd import scene as sc

    hg = download_grid(ds.grid)
    he, keep = download_env(ds.env)
    p = sc.fill_params(cfg, hg.grid, he, ds.camera)
    host_nrc.fill(p.nrc)
    dp = r.params()
    p.nrc.nrc_buffer_index = dp.nrc.nrc_buffer_index
    p.nrc.offset_to_select_unbiased_path = dp.nrc.offset_to_select_unbiased_path
    p.nrc.radiance_scale = dp.nrc.radiance_scale
    p.nrc.path_termination_factor = dp.nrc.path_termination_factor
    p.rng_states, p.accum = ptr(rng), ptr(accum)
    return p, (hg, he, keep)


@pytest.mark.parametrize("tracking", [0, 1])
def test_training_data_path_bit_exact(tracking):
    """prepare -> GENERATE_TRAINING_DATA trace -> propagate -> shuffle, GPU vs oracle on the same inputs."""
    _training_data_case(tracking=tracking)


@pytest.mark.parametrize("width,height,lo,hi", [(512, 384, 1, 4), (1024, 768, 4, 1 << 20)])
def test_training_data_overflow_bit_exact(width, height, lo, hi):
    """More training records than the buffer holds (long paths: a thick medium, a large path-termination
    factor and Russian roulette only after 64 vertices).  Paths end at their first record past the
    buffer, rebuilt from the staged records; with more records than the staging area holds (the second
    case) some pixels are traced again.  The results are still the oracle's."""
    from rtcamp8_amd import _abi

    num = _training_data_case(width=width, height=height, termination=1.0e4, rr_after=64, density=400.0)
    print(f"[training data] {num} records", flush=True)
    assert lo * _abi.TRAIN_BUFFER_SIZE < num <= hi * _abi.TRAIN_BUFFER_SIZE


def _training_data_case(width=96, height=64, tracking=0, termination=None, rr_after=None, density=None):
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call

    print(f"[training data] {width}x{height} tracking {tracking} termination {termination}", flush=True)
    cfg, ds, r = _nrc_case(width=width, height=height, tracking=tracking, density=density)
    b = r.buffers
    r.buf_idx = 0
    p = r.params()
    p.nrc.offset_to_select_unbiased_path = 5
    if termination is not None:
        p.nrc.path_termination_factor = termination
    if rr_after is not None:
        p.rr_after_path_length = rr_after
    # snapshot of the inputs, then the GPU steps
    before = HostNrc(b)
    rng0 = r.rng.download(np.uint64, cfg.width * cfg.height)
    call("rtc8x_nrc_prepare", None, C.byref(p), 1)
    call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_GENERATE_TRAINING_DATA, 1)
    ts = b.train_image_size.download(np.int32, 4)
    n_term = int(ts[0] * ts[1])
    inferred = np.random.default_rng(19).random(b.inferred_radiance.nbytes // 4).astype(np.float32) * 2 - 0.2
    b.inferred_radiance.upload(inferred)
    after_trace = HostNrc(b)
    call("rtc8x_nrc_propagate", None, C.byref(p))
    after_prop = HostNrc(b)
    call("rtc8x_nrc_shuffle", None, C.byref(p))
    after_shuf = HostNrc(b)
    gpu_rng = r.rng.download(np.uint64, cfg.width * cfg.height)

    # oracle on host copies
    rng = rng0.copy()
    accum = np.zeros(3 * cfg.width * cfg.height)
    hp, keep = _host_params(cfg, ds, r, before, rng, accum)
    hp.nrc.nrc_buffer_index = 0
    hp.nrc.offset_to_select_unbiased_path = 5
    hp.nrc.path_termination_factor = p.nrc.path_termination_factor
    hp.rr_after_path_length = p.rr_after_path_length
    oracle().oracle_nrc_prepare(C.byref(hp), 1)
    assert oracle().oracle_trace_f32(C.byref(hp), _abi.ENTRY_GENERATE_TRAINING_DATA, 1, 0, 10 ** 6, 4) == 0
    a, g = before.a, after_trace.a
    num = int(a["num_training_data"][0])
    assert num == int(g["num_training_data"][0]) and num > 100
    assert np.array_equal(a["train_image_size"], g["train_image_size"])
    assert np.array_equal(rng, gpu_rng)
    k = min(num, _abi.TRAIN_BUFFER_SIZE)
    assert np.array_equal(a["train_queries0"][:14 * k].view(np.uint32), g["train_queries0"][:14 * k].view(np.uint32))
    assert np.array_equal(a["train_targets0"][:3 * k].view(np.uint32), g["train_targets0"][:3 * k].view(np.uint32))
    assert np.array_equal(a["train_vertex_info"][:6 * k], g["train_vertex_info"][:6 * k])
    assert np.array_equal(a["train_suffix_terminals"][:4 * n_term], g["train_suffix_terminals"][:4 * n_term])
    has_q = a["train_suffix_terminals"][:4 * n_term].reshape(-1, 4)[:, 1] == 1
    qi = a["inference_queries"].reshape(-1, 14)[:n_term][has_q]
    qg = g["inference_queries"].reshape(-1, 14)[:n_term][has_q]
    assert np.array_equal(qi.view(np.uint32), qg.view(np.uint32))
    # propagate and shuffle from the GPU's post-trace state
    a2 = HostNrc(b)
    a2.a = {k2: v.copy() for k2, v in after_trace.a.items()}
    hp2, keep2 = _host_params(cfg, ds, r, a2, rng, accum)
    hp2.nrc.nrc_buffer_index = 0
    oracle().oracle_nrc_propagate(C.byref(hp2))
    assert np.array_equal(a2.a["train_targets0"].view(np.uint32), after_prop.a["train_targets0"].view(np.uint32))
    oracle().oracle_nrc_shuffle(C.byref(hp2))
    for key in ("train_queries1", "train_targets1", "data_shufflers", "target_min_max", "target_avg"):
        assert np.array_equal(a2.a[key].view(np.uint32), after_shuf.a[key].view(np.uint32)), key
    r.close()
    ds.close()
    return num


@pytest.mark.parametrize("ordered", [False, True])
def test_render_with_nrc_bit_exact(ordered):
    """PATH_TRACE_WITH_NRC trace -> (synthetic) inference -> accumulate, GPU vs oracle; optionally with
    lanes taking the pixels in a random work order (results must not depend on it)."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    cfg, ds, r = _nrc_case()
    b = r.buffers
    p = r.params()
    if ordered:
        perm = np.random.default_rng(21).permutation(cfg.width * cfg.height).astype(np.uint32)
        dperm = DeviceBuffer.from_array(perm)
        p.work_order = dperm.ptr
        p.pixel_cost = r.cost.ptr
    rng0 = r.rng.download(np.uint64, cfg.width * cfg.height)
    before = HostNrc(b)
    call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1)
    inferred = np.random.default_rng(20).random(b.inferred_radiance.nbytes // 4).astype(np.float32) * 3 - 0.5
    b.inferred_radiance.upload(inferred)
    call("rtc8x_nrc_accumulate", None, C.byref(p))
    after = HostNrc(b)
    gpu_accum = r.accum.download(np.float64, 3 * cfg.width * cfg.height)
    gpu_rng = r.rng.download(np.uint64, cfg.width * cfg.height)

    rng = rng0.copy()
    accum = np.zeros(3 * cfg.width * cfg.height)
    hp, keep = _host_params(cfg, ds, r, before, rng, accum)
    assert oracle().oracle_trace_f32(C.byref(hp), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1, 0, 10 ** 6, 4) == 0
    before.a["inferred_radiance"][:] = inferred
    oracle().oracle_nrc_accumulate(C.byref(hp))
    n = cfg.width * cfg.height
    a, g = before.a, after.a
    assert np.array_equal(rng, gpu_rng)
    assert np.array_equal(a["inference_terminals"].view(np.uint32), g["inference_terminals"].view(np.uint32))
    assert np.array_equal(a["per_frame_contribution"].view(np.uint32), g["per_frame_contribution"].view(np.uint32))
    has_q = (a["inference_terminals"].reshape(-1, 4)[:n, 3] & 1) == 1
    assert has_q.sum() > 10
    assert np.array_equal(a["inference_queries"].reshape(-1, 14)[:n][has_q].view(np.uint32),
                          g["inference_queries"].reshape(-1, 14)[:n][has_q].view(np.uint32))
    assert np.array_equal(accum.view(np.uint64), gpu_accum.view(np.uint64))
    r.close()
    ds.close()


def test_render_with_nrc_several_samples_per_launch():
    """A PATH_TRACE_WITH_NRC launch of 3 samples (slices k * image of the render buffers) matches the
    oracle bit for bit, and its accumulated sums equal those of three one-sample launches."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call

    cfg, ds, r = _nrc_case()
    assert r.samples_per_launch >= 3
    b = r.buffers
    n = cfg.width * cfg.height
    nq = ((n + 127) // 128) * 128
    p = r.params()
    p.nrc.num_samples = 3
    rng0 = r.rng.download(np.uint64, n)
    before = HostNrc(b)
    call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, 3)
    inferred = np.random.default_rng(22).random(b.inferred_radiance.nbytes // 4).astype(np.float32) * 3 - 0.5
    b.inferred_radiance.upload(inferred)
    call("rtc8x_nrc_accumulate", None, C.byref(p))
    after = HostNrc(b)
    gpu_accum = r.accum.download(np.float64, 3 * n)
    gpu_rng = r.rng.download(np.uint64, n)
    with pytest.raises(Exception):  # the launch must trace exactly nrc.num_samples samples
        call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, 2)

    rng = rng0.copy()
    accum = np.zeros(3 * n)
    hp, keep = _host_params(cfg, ds, r, before, rng, accum)
    hp.nrc.num_samples = 3
    assert oracle().oracle_trace_f32(C.byref(hp), _abi.ENTRY_PATH_TRACE_WITH_NRC, 3, 0, 10 ** 6, 4) == 0
    before.a["inferred_radiance"][:] = inferred
    oracle().oracle_nrc_accumulate(C.byref(hp))
    a, g = before.a, after.a
    assert np.array_equal(rng, gpu_rng)
    term_a = a["inference_terminals"].reshape(-1, 4)[:3 * n]
    term_g = g["inference_terminals"].reshape(-1, 4)[:3 * n]
    assert np.array_equal(term_a.view(np.uint32), term_g.view(np.uint32))
    assert np.array_equal(a["per_frame_contribution"][:9 * n].view(np.uint32),
                          g["per_frame_contribution"][:9 * n].view(np.uint32))
    qa = a["inference_queries"].reshape(-1, 14)
    qg = g["inference_queries"].reshape(-1, 14)
    for k in range(3):
        has_q = (term_a[k * n:(k + 1) * n, 3] & 1) == 1
        assert has_q.sum() > 10
        sl = slice(k * nq, k * nq + n)
        assert np.array_equal(qa[sl][has_q].view(np.uint32), qg[sl][has_q].view(np.uint32))
    assert np.array_equal(accum.view(np.uint64), gpu_accum.view(np.uint64))

    # three one-sample launches from the same state give the same sums
    r.rng.upload(rng0)
    r.accum.fill(0)
    for k in range(3):
        p1 = r.params()
        p1.nrc.num_samples = 1
        call("rtc8x_trace", None, C.byref(p1), _abi.ENTRY_PATH_TRACE_WITH_NRC, 1)
        b.inferred_radiance.upload(inferred[3 * k * nq: 3 * (k + 1) * nq].copy())
        call("rtc8x_nrc_accumulate", None, C.byref(p1))
    assert np.array_equal(r.accum.download(np.float64, 3 * n).view(np.uint64), gpu_accum.view(np.uint64))
    assert np.array_equal(r.rng.download(np.uint64, n), gpu_rng)
    r.close()
    ds.close()


@pytest.mark.parametrize("precision", ["fp16", "bf16_fp8"])
def test_render_inference_at_query_slots_matches_the_full_batch(precision):
    """rtc8x_nrc_infer_render (two samples per launch, dense 128^3 cloud at 96 x 64: about a fifth of the
    slots hold a query) evaluates the slots with a query and leaves the others as they are (a byte
    pattern written before the call).  At those slots it equals rtc8x_nrc_infer over the whole batch: bit
    for bit in FP16;
    with E4M3 inference, whose per-tile activation scales then come from the queries alone, within
    0.05 relative RMS."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.device import DeviceBuffer
    from rtcamp8_amd.nrc import NrcRenderer, round_up

    prec = {"fp16": _abi.NRC_FP16, "bf16_fp8": _abi.NRC_BF16_FP8}[precision]
    cfg = sc.CONFIGS["dense128"].with_(width=96, height=64)
    print(f"[infer render] {precision} {cfg.width}x{cfg.height}", flush=True)
    ds = sc.DeviceScene(cfg)
    r = NrcRenderer(ds, num_hidden_layers=2, precision=prec, encoding=_abi.NRC_HASH_GRID, samples_per_launch=2)
    full = None
    try:
        r.reset(7)
        r.render_frame(spp=2)
        r.image()
        b = r.buffers
        n, q = cfg.width * cfg.height, round_up(cfg.width * cfg.height, 128)
        b.inferred_radiance.fill(0xA5)
        p = r.params()
        p.nrc.num_samples = 2
        r.cache.infer_render(None, p)
        got = b.inferred_radiance.download(np.float32, 3 * 2 * q).reshape(-1, 3)
        full = DeviceBuffer(12 * 2 * q)
        r.cache.infer(None, b.inference_queries.ptr, 2 * q, full.ptr)
        want = full.download(np.float32, 3 * 2 * q).reshape(-1, 3)
        has = (b.inference_terminals.download(np.uint32, 4 * 2 * n).reshape(2, n, 4)[..., 3] & 1) == 1
    finally:
        if full is not None:
            full.free()
        r.close()
        ds.close()
    slots = (np.arange(2)[:, None] * q + np.arange(n)[None, :])
    live, dead = slots[has], np.setdiff1d(np.arange(2 * q), slots[has])
    print(f"[infer render] {has.mean():.3f} of the slots hold a query", flush=True)
    assert 0.05 < has.mean() < 0.95
    assert np.all(got[dead].view(np.uint32) == 0xA5A5A5A5)
    if precision == "fp16":
        assert np.array_equal(got[live].view(np.uint32), want[live].view(np.uint32))
    else:
        rel = np.sqrt(np.mean((got[live] - want[live]) ** 2) / np.mean(want[live] ** 2))
        print(f"[infer render] E4M3: relative RMS difference {rel:.4f}", flush=True)
        assert rel < 0.05
