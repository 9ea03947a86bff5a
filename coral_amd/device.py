"""Device memory, streams and event timing through libcoralhip's own utilities.

No PyTorch: device buffers are plain HIP allocations owned by ``DeviceArray``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import call, load


class DeviceArray:
    """A flat device buffer with a numpy dtype and element count."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = ctypes.c_void_p()
        call("coral_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if d.nbytes:
            call("coral_memcpy_htod", d.ptr, a.ctypes.data, d.nbytes)
        return d

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.nbytes:
            call("coral_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        call("coral_memcpy_htod", self.ptr, a.ctypes.data, self.nbytes)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            load().coral_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    call("coral_synchronize")


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    call("coral_device_name", buf, 256)
    return buf.value.decode()


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        call("coral_stream_create", ctypes.byref(s))
        self.handle = s.value

    def synchronize(self):
        call("coral_stream_synchronize", self.handle)

    def destroy(self):
        if self.handle:
            call("coral_stream_destroy", self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        call("coral_event_create", ctypes.byref(e))
        self.handle = e.value

    def record(self, stream=None):
        call("coral_event_record", self.handle, stream.handle if stream else None)

    def synchronize(self):
        call("coral_event_synchronize", self.handle)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        call("coral_event_elapsed_ms", ctypes.byref(ms), self.handle, end.handle)
        return ms.value

    def __del__(self):
        try:
            load().coral_event_destroy(self.handle)
        except Exception:
            pass


def fill_random(arr: DeviceArray, seed: int, kind: str = "uniform", stream=None) -> None:
    """Seeded synthetic data on the device (kind: 'uniform' in [-1,1) or 'normal')."""
    k = {"uniform": 0, "normal": 1}[kind]
    count = arr.count
    if arr.dtype == np.float32:
        call("coral_fill_random_f32", arr.ptr, count, seed, k, stream.handle if stream else None)
    elif arr.dtype == np.float64:
        call("coral_fill_random_f64", arr.ptr, count, seed, k, stream.handle if stream else None)
    elif arr.dtype == np.complex64:
        call("coral_fill_random_f32", arr.ptr, 2 * count, seed, k, stream.handle if stream else None)
    elif arr.dtype == np.complex128:
        call("coral_fill_random_f64", arr.ptr, 2 * count, seed, k, stream.handle if stream else None)
    else:
        raise TypeError(arr.dtype)
