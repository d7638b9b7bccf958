"""Several instances in one persistent launch (rtc8x_trace_batch): every instance's results are those of
its own rtc8x_trace launch, bit for bit, whatever the shares of the device."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes():
    from rtcamp8_amd import scene as sc

    # uneven instances: different volumes (dense and sparse, 64^3 to 256^3), densities, images and spp
    return [
        sc.CONFIGS["dense128"].with_(width=96, height=64, fbm=sc.FbmConfig(resolution=64, seed=0x5EED0001)),
        sc.CONFIGS["sparse512"].with_(width=160, height=90, fbm=sc.FbmConfig(resolution=256, seed=0x5EED0002, sparse=True,
                                                                               active_leaf_fraction=0.25, contrast=4.0)),
        sc.CONFIGS["dense128"].with_(width=48, height=128, density_coeff=5.0,
                                     fbm=sc.FbmConfig(resolution=128, seed=0x5EED0003)),
    ]


@pytest.mark.parametrize("shares", [None, [1.0, 6.0, 0.5]])
def test_plain_batch_matches_single_launches(shares):
    from rtcamp8_amd import batch
    from rtcamp8_amd.scene import DeviceScene
    from rtcamp8_amd.tracer import PathTracer

    cfgs = _scenes()
    spp = [3, 2, 5]
    print(f"[batch] plain, shares {shares}", flush=True)
    scenes = [DeviceScene(c) for c in cfgs]
    single = [PathTracer(s) for s in scenes]
    batched = [PathTracer(s) for s in scenes]
    for t, k in zip(single, spp):
        t.trace(k)
    batch.trace_plain(batched, spp, shares)
    for i, (a, b, c) in enumerate(zip(single, batched, cfgs)):
        n = c.width * c.height
        assert np.array_equal(a.accum.download(np.float64, 3 * n), b.accum.download(np.float64, 3 * n)), i
        assert np.array_equal(a.rng.download(np.uint64, n), b.rng.download(np.uint64, n)), i
        assert np.array_equal(a.stats.download(np.uint64, 4), b.stats.download(np.uint64, 4)), i
        assert a.image().mean() > 0
    # two more launches, on the cost-ordered work lists the first one recorded
    for t, k in zip(single, spp):
        t.trace(k)
        t.trace(k)
    batch.trace_plain(batched, spp, shares)
    batch.trace_plain(batched, spp, shares)
    for i, (a, b, c) in enumerate(zip(single, batched, cfgs)):
        n = c.width * c.height
        assert np.array_equal(a.accum.download(np.float64, 3 * n), b.accum.download(np.float64, 3 * n)), i
    for t in single + batched:
        t.close()
    for s in scenes:
        s.close()


def test_nrc_batch_matches_single_launches():
    from rtcamp8_amd import _abi, batch
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.nrc import NrcRenderer
    from rtcamp8_amd.scene import DeviceScene

    cfgs = _scenes()[:2]
    k = [2, 1]
    scenes = [DeviceScene(c) for c in cfgs]
    rs = [NrcRenderer(s, num_hidden_layers=2, samples_per_launch=2) for s in scenes]
    keys = ("inference_queries", "inference_terminals", "per_frame_contribution")

    def outputs(r):
        return [getattr(r.buffers, key).download(np.uint32, getattr(r.buffers, key).nbytes // 4) for key in keys]

    rng0 = [r.rng.download(np.uint64, c.width * c.height) for r, c in zip(rs, cfgs)]
    want = []
    for r, kk in zip(rs, k):
        p = r.params()
        p.nrc.num_samples = kk
        call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, kk)
        want.append((outputs(r), r.rng.download(np.uint64, r.width * r.height)))
    for r, g in zip(rs, rng0):
        r.rng.upload(g)
        for key in keys:
            getattr(r.buffers, key).fill(0)
    params = []
    for r, kk in zip(rs, k):
        p = r.params()
        p.nrc.num_samples = kk
        params.append(p)
    batch._launch(params, k, _abi.ENTRY_PATH_TRACE_WITH_NRC, None, None)
    for i, r in enumerate(rs):
        got = outputs(r)
        for key, a, b in zip(keys, want[i][0], got):
            assert np.array_equal(a, b), (i, key)
        assert np.array_equal(want[i][1], r.rng.download(np.uint64, r.width * r.height)), i
    # the batched render step runs end to end
    batch.render_nrc(rs, k)
    for r in rs:
        assert np.all(np.isfinite(r.image()))
        r.close()
    for s in scenes:
        s.close()


def test_batch_rejects_bad_input():
    from rtcamp8_amd import _abi, batch
    from rtcamp8_amd._lib import Rtc8xError
    from rtcamp8_amd.scene import DeviceScene
    from rtcamp8_amd.tracer import PathTracer

    s = DeviceScene(_scenes()[0])
    t = PathTracer(s)
    p = t.params()
    with pytest.raises(Rtc8xError):
        batch._launch([p] * 17, [1] * 17, _abi.ENTRY_PATH_TRACE, None, None)  # more than 16 instances
    with pytest.raises(Rtc8xError):
        batch._launch([p], [1], _abi.ENTRY_GENERATE_TRAINING_DATA, None, None)
    with pytest.raises(Rtc8xError):
        batch._launch([p, p], [1, 1], _abi.ENTRY_PATH_TRACE, [1.0, 0.0], None)
    bad = t.params()
    bad.accum = None
    with pytest.raises(Rtc8xError, match="instance 1"):
        batch._launch([p, bad], [1, 1], _abi.ENTRY_PATH_TRACE, None, None)
    t.close()
    s.close()


@pytest.mark.parametrize("entry", ["plain", "nrc"])
def test_scheduling_knobs_do_not_change_results(entry):
    """Event threshold and tile shape change the schedule only: accumulators, PCG states, statistics and NRC
    trace outputs are those of the default schedule, bit for bit."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.nrc import NrcRenderer
    from rtcamp8_amd.scene import DeviceScene
    from rtcamp8_amd.tracer import PathTracer

    cfg = _scenes()[1]
    s = DeviceScene(cfg)
    n = cfg.width * cfg.height
    outs = []
    for knobs in ((0, 0), (2, 1), (5, 3), (8, 2)):
        if entry == "plain":
            t = PathTracer(s, schedule=False)
            t.knobs = knobs
            t.trace(3)
            outs.append((t.accum.download(np.float64, 3 * n), t.rng.download(np.uint64, n), t.stats.download(np.uint64, 4)))
            t.close()
        else:
            r = NrcRenderer(s, num_hidden_layers=2, samples_per_launch=2)
            r.knobs = knobs
            p = r.params()
            p.nrc.num_samples = 2
            call("rtc8x_trace", None, C.byref(p), _abi.ENTRY_PATH_TRACE_WITH_NRC, 2)
            b = r.buffers
            outs.append(tuple(getattr(b, k).download(np.uint32, getattr(b, k).nbytes // 4)
                              for k in ("inference_queries", "inference_terminals", "per_frame_contribution")) +
                        (r.rng.download(np.uint64, n),))
            r.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    s.close()
