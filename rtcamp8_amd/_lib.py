"""Loads the in-tree gfx950 library (rtcamp8_amd/librtc8x.so) and binds every C-ABI symbol.

There is no CPU fallback: if the library is missing, or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import _abi

LIB_PATH = Path(__file__).resolve().parent / "librtc8x.so"


class Rtc8xError(RuntimeError):
    """A C-ABI call returned a non-zero rtc8x_status."""

    def __init__(self, func: str, status: int, message: str):
        super().__init__(f"{func} failed with status {status}: {message}")
        self.func = func
        self.status = status


_lib = None


def lib() -> C.CDLL:
    """The loaded library with argtypes/restype set for every declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the gfx950 library is required, there is no CPU fallback)")
    handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_GLOBAL)
    for name, (restype, argtypes) in _abi.PROTOTYPES.items():
        fn = getattr(handle, name)  # AttributeError if a declared symbol is not exported
        fn.restype = restype
        fn.argtypes = argtypes
    if handle.rtc8x_abi_version() != 1:
        raise RuntimeError("librtc8x ABI version mismatch")
    for name, cls in _abi.STRUCTS.items():
        size = handle.rtc8x_sizeof(name.encode())
        if size != C.sizeof(cls):
            raise RuntimeError(f"ctypes layout of {name} is {C.sizeof(cls)} bytes, the library's is {size}")
    _lib = handle
    return _lib


def check(func: str, status: int) -> None:
    if status != _abi.RTC8X_OK:
        msg = lib().rtc8x_last_error()
        raise Rtc8xError(func, status, msg.decode() if msg else "")


def call(name: str, *args):
    """Calls an rtc8x_* function that returns rtc8x_status and raises on failure."""
    fn = getattr(lib(), name)
    check(name, fn(*args))
