"""rtcamp8_amd -- MI355X (gfx950) volumetric path tracer with a neural radiance cache.

The compute path is the in-tree HIP library ``librtc8x.so`` (C ABI declared in include/rtc8x.h);
this package is the host-side mirror of the reference's renderer interface around it.
"""
from ._lib import LIB_PATH, Rtc8xError, lib  # noqa: F401
from . import _abi  # noqa: F401

__all__ = ["lib", "LIB_PATH", "Rtc8xError", "_abi"]
