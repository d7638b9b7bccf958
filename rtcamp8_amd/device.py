"""Device memory, streams and timers over the C ABI (the role cudau::Buffer / cudau::Timer play in
the reference, common/utils/cuda_util.h:175-300)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib


def device_count() -> int:
    n = C.c_int32(0)
    call("rtc8x_device_count", C.byref(n))
    return int(n.value)


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("rtc8x_device_name", buf, 256)
    return buf.value.decode()


def device_cus() -> int:
    """Compute units of the current device (rtc8x_device_name reports them as "..., N CUs)")."""
    import re

    m = re.search(r"(\d+) CUs\)", device_name())
    if not m:
        raise RuntimeError(f"no CU count in the device name {device_name()!r}")
    return int(m.group(1))


def synchronize() -> None:
    call("rtc8x_device_synchronize")


def wait_stream(stream) -> None:
    """Waits for the work queued on `stream` (None: the null stream, which the blocking copies of
    DeviceBuffer already wait for).  Streams from Stream() do not block on the null stream, so a copy
    on it does not wait for them: read their results after this."""
    if stream:
        call("rtc8x_stream_synchronize", C.c_void_p(stream))


class DeviceBuffer:
    """An owned device allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("rtc8x_malloc", C.byref(p), self.nbytes)
        self.ptr = int(p.value or 0)

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(max(arr.nbytes, 1))
        buf.upload(arr)
        return buf

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        buf = cls(max(int(nbytes), 1))
        buf.fill(0)
        return buf

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes, "upload out of range"
        call("rtc8x_memcpy_htod", C.c_void_p(self.ptr + offset), arr.ctypes.data_as(C.c_void_p), arr.nbytes)

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes, "download out of range"
        call("rtc8x_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset), out.nbytes)
        return out

    def fill(self, byte_value: int) -> None:
        call("rtc8x_memset", C.c_void_p(self.ptr), int(byte_value), self.nbytes)

    def free(self) -> None:
        if self.ptr:
            call("rtc8x_free", C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr and lib() is not None:
                lib().rtc8x_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


def download_ptr(ptr: int, dtype, count: int) -> np.ndarray:
    dtype = np.dtype(dtype)
    out = np.empty(count, dtype=dtype)
    if count:
        call("rtc8x_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes)
    return out


class Stream:
    def __init__(self):
        s = C.c_void_p()
        call("rtc8x_stream_create", C.byref(s))
        self.handle = s.value

    def synchronize(self) -> None:
        call("rtc8x_stream_synchronize", C.c_void_p(self.handle))

    def destroy(self) -> None:
        if self.handle:
            call("rtc8x_stream_destroy", C.c_void_p(self.handle))
            self.handle = None


class Event:
    """A HIP event for ordering streams: record() marks the work queued so far on a stream, and
    make_wait() makes later work on another stream wait for it."""

    def __init__(self):
        e = C.c_void_p()
        call("rtc8x_event_create", C.byref(e))
        self.handle = e.value

    def record(self, stream) -> None:
        call("rtc8x_event_record", C.c_void_p(self.handle), C.c_void_p(stream))

    def make_wait(self, stream) -> None:
        call("rtc8x_stream_wait_event", C.c_void_p(stream), C.c_void_p(self.handle))

    def destroy(self) -> None:
        if self.handle:
            call("rtc8x_event_destroy", C.c_void_p(self.handle))
            self.handle = None


class EventTimer:
    """A pair of HIP events recorded on a stream (cudau::Timer, common/utils/cuda_util.h:257-300)."""

    def __init__(self):
        a, b = C.c_void_p(), C.c_void_p()
        call("rtc8x_event_create", C.byref(a))
        call("rtc8x_event_create", C.byref(b))
        self.start_ev, self.stop_ev = a.value, b.value

    def start(self, stream) -> None:
        call("rtc8x_event_record", C.c_void_p(self.start_ev), C.c_void_p(stream))

    def stop(self, stream) -> None:
        call("rtc8x_event_record", C.c_void_p(self.stop_ev), C.c_void_p(stream))

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        call("rtc8x_event_elapsed_ms", C.byref(ms), C.c_void_p(self.start_ev), C.c_void_p(self.stop_ev))
        return float(ms.value)

    def destroy(self) -> None:
        for ev in (self.start_ev, self.stop_ev):
            if ev:
                call("rtc8x_event_destroy", C.c_void_p(ev))
        self.start_ev = self.stop_ev = None
