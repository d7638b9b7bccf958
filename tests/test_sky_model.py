"""The Hosek-Wilkie state for any sun on the host (no GPU): rtc8x_hosek_bands against the vendored model's
arhosekskymodelstate_alloc_init (tests/golden/hosek_states.npz, tests/golden/make_sky_model.py), bit for
bit after the reference's float casts, over elevations 0 .. pi/2 and turbidities 1 .. 10 (integer levels,
fractions and the clamped top level)."""
import numpy as np
import pytest
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden" / "hosek_states.npz"


def test_band_states_match_the_model():
    from rtcamp8_amd.sky import HosekModel

    g = np.load(GOLDEN)
    m = HosekModel()
    try:
        for k in range(len(g["elevation"])):
            got = m.bands(float(g["elevation"][k]), float(g["turbidity"][k]))
            assert np.array_equal(got.view(np.uint32), g["states"][k].view(np.uint32)), k
    finally:
        m.close()


def test_out_of_range_suns_are_rejected():
    from rtcamp8_amd._lib import Rtc8xError
    from rtcamp8_amd.sky import HosekModel

    m = HosekModel()
    try:
        for el, t in ((0.3, 0.5), (0.3, 10.5), (-0.1, 3.0), (1.6, 3.0)):
            with pytest.raises(Rtc8xError):
                m.bands(el, t)
    finally:
        m.close()


def test_model_image_is_validated():
    import ctypes as C

    from rtcamp8_amd._lib import Rtc8xError, call

    h = C.c_void_p()
    with pytest.raises(Rtc8xError):
        call("rtc8x_hosek_model_create", b"RTC8XHW1" + b"\0" * 100, 108, C.byref(h))
