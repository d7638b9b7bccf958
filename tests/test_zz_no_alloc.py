"""rtc8x_trace_batch_static, the batched trace that allocates nothing (its instance table is a fixed device
array): on two calls in a row, the second with fewer and smaller instances, every instance's accumulator,
PCG states and statistics equal those of rtc8x_trace_batch on the same inputs, bit for bit; more instances
than the table holds is an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfgs(big: bool):
    from rtcamp8_amd import scene as sc

    if big:
        return [
            sc.CONFIGS["dense128"].with_(width=96, height=64, fbm=sc.FbmConfig(resolution=64, seed=0x5EED0001)),
            sc.CONFIGS["sparse512"].with_(width=160, height=90, sky_width=256, sky_height=128,
                                          fbm=sc.FbmConfig(resolution=256, seed=0x5EED0002, sparse=True)),
            sc.CONFIGS["dense128"].with_(width=48, height=128, density_coeff=5.0,
                                         fbm=sc.FbmConfig(resolution=128, seed=0x5EED0003)),
        ], [2, 1, 3]
    return [
        sc.CONFIGS["dense128"].with_(width=40, height=24, fbm=sc.FbmConfig(resolution=64, seed=0x5EED0004)),
        sc.CONFIGS["dense128"].with_(width=24, height=40, density_coeff=8.0,
                                     fbm=sc.FbmConfig(resolution=64, seed=0x5EED0005)),
    ], [1, 2]


def _batch(name, tracers, spp, shares=None):
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call

    n = len(tracers)
    params = (_abi.FrameParams * n)(*[t.params() for t in tracers])
    ns = (C.c_uint32 * n)(*spp)
    sh = None if shares is None else (C.c_float * n)(*shares)
    call(name, None, params, ns, sh, n, _abi.ENTRY_PATH_TRACE)


def test_static_batch_matches_the_allocating_batch_on_two_calls():
    from rtcamp8_amd.scene import DeviceScene
    from rtcamp8_amd.tracer import PathTracer

    for big, shares in ((True, [1.0, 4.0, 0.5]), (False, None)):
        cfgs, spp = _cfgs(big)
        print(f"[no-alloc] {len(cfgs)} instances, spp {spp}", flush=True)
        scenes = [DeviceScene(c) for c in cfgs]
        ref = [PathTracer(s, schedule=False) for s in scenes]
        got = [PathTracer(s, schedule=False) for s in scenes]
        _batch("rtc8x_trace_batch", ref, spp, shares)
        _batch("rtc8x_trace_batch_static", got, spp, shares)
        for i, (a, b, c) in enumerate(zip(ref, got, cfgs)):
            n = c.width * c.height
            assert np.array_equal(a.accum.download(np.float64, 3 * n), b.accum.download(np.float64, 3 * n)), i
            assert np.array_equal(a.rng.download(np.uint64, n), b.rng.download(np.uint64, n)), i
            assert np.array_equal(a.stats.download(np.uint64, 4), b.stats.download(np.uint64, 4)), i
            assert b.stats.download(np.uint64, 4)[0] == n * spp[i]
        for t in ref + got:
            t.close()
        for s in scenes:
            s.close()


def test_static_batch_rejects_more_instances_than_its_table():
    from rtcamp8_amd._lib import Rtc8xError
    from rtcamp8_amd.scene import DeviceScene
    from rtcamp8_amd.tracer import PathTracer

    cfgs, _ = _cfgs(False)
    s = DeviceScene(cfgs[0])
    t = PathTracer(s, schedule=False)
    with pytest.raises(Rtc8xError, match="fixed instance table holds 16"):
        _batch("rtc8x_trace_batch_static", [t] * 17, [1] * 17)
    t.close()
    s.close()
