"""The march probe (rtc8x_march_probe), the latency ceiling the bench compares the tracer with: both of its
walks run to their lookup budget, count paths, and replay the tracker they stand for (the local walk does
the local tracker's few lookups per path, the global one the global tracker's many)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _probe(tracking: int, budget: int = 48, waves: int = 3):
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer, device_cus, synchronize
    from rtcamp8_amd.tracer import PathTracer

    cfg = sc.CONFIGS["sparse512"].with_(
        width=160, height=96, spp=2, tracking=tracking,
        fbm=sc.FbmConfig(resolution=256, seed=0x5EED0002, sparse=True, active_leaf_fraction=0.25))
    print(f"[probe] tracking {tracking}, budget {budget}, {waves} waves/SIMD", flush=True)
    scene = sc.DeviceScene(cfg)
    pt = PathTracer(scene)
    pt.trace(cfg.spp)  # the tracer's own lookups per path on this frame
    c = pt.counters()
    params = pt.params()
    light = (C.c_float * 3)(*[float(v) for v in sc.sun_direction(cfg)])
    stats = DeviceBuffer.zeros(16)
    call("rtc8x_march_probe", None, C.byref(params), light, cfg.spp, budget, waves, C.c_void_p(stats.ptr))
    synchronize()
    looked, paths = (int(v) for v in stats.download(np.uint64, 2))
    stats.free()
    pt.close()
    scene.close()
    # a lane stops once it has made `budget` lookups; a speculative pair can take it one past
    return looked, paths, c["lookups"] / max(c["paths"], 1), device_cus() * waves * 256 * (budget + 1)


def test_march_probe_walks_both_trackers():
    res = {}
    for tracking in (0, 1):
        looked, paths, trace_lpp, cap = _probe(tracking)
        assert 0 < looked <= cap, (tracking, looked, cap)
        assert paths > 0, tracking
        # most lanes spend their whole budget: the launch has no tail
        assert looked > 0.5 * cap, (tracking, looked, cap)
        res[tracking] = (looked / paths, trace_lpp)
    (g_probe, g_trace), (l_probe, l_trace) = res[0], res[1]
    # the probe replays the tracker's walk: lookups per path of the same order as the tracer's, and far
    # fewer for the local walk, as for the tracer
    assert l_trace < 0.5 * g_trace, res
    assert l_probe < 0.5 * g_probe, res
    for probe, trace in ((g_probe, g_trace), (l_probe, l_trace)):
        assert 0.25 * trace < probe < 4.0 * trace, res


def test_march_probe_rejects_occupancy_beyond_its_register_budget():
    """At most 4 waves per SIMD: the probe's registers are allotted for 3 (the tracer's occupancy) or 4."""
    from rtcamp8_amd._lib import Rtc8xError

    for tracking in (0, 1):
        with pytest.raises(Rtc8xError):
            _probe(tracking, waves=8)
