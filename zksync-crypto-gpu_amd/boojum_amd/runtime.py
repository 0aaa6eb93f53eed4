"""era_cudart mirror (crates/cudart/src): device memory, streams and events over the C-ABI."""
import ctypes

import numpy as np

from ._lib import check, lib

H2D, D2H, D2D = 1, 2, 3


def device_count():
    n = ctypes.c_int(0)
    rc = lib().ba_get_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def set_device(device):
    check(lib().ba_set_device(device), "set_device")


def synchronize():
    check(lib().ba_device_synchronize(), "device_synchronize")


def mem_get_info():
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().ba_mem_get_info(ctypes.byref(free), ctypes.byref(total)), "mem_get_info")
    return free.value, total.value


class CudaStream:
    """Non-blocking stream at the default priority (crates/cudart/src/stream.rs:17)."""

    def __init__(self, default=False):
        self.handle = ctypes.c_void_p(None)
        self._owned = not default
        if not default:
            check(lib().ba_stream_create(ctypes.byref(self.handle)), "stream_create")

    @classmethod
    def default(cls):
        return cls(default=True)

    def synchronize(self):
        check(lib().ba_stream_synchronize(self.handle), "stream_synchronize")

    def wait_event(self, event):
        check(lib().ba_stream_wait_event(self.handle, event.handle), "stream_wait_event")

    def destroy(self):
        if self._owned and self.handle.value:
            check(lib().ba_stream_destroy(self.handle), "stream_destroy")
            self.handle = ctypes.c_void_p(None)


class CudaEvent:
    def __init__(self):
        self.handle = ctypes.c_void_p(None)
        check(lib().ba_event_create(ctypes.byref(self.handle)), "event_create")

    def record(self, stream):
        check(lib().ba_event_record(self.handle, stream.handle), "event_record")

    def synchronize(self):
        check(lib().ba_event_synchronize(self.handle), "event_synchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        check(lib().ba_event_elapsed_ms(ctypes.byref(ms), self.handle, end.handle), "event_elapsed")
        return ms.value

    def destroy(self):
        if self.handle.value:
            check(lib().ba_event_destroy(self.handle), "event_destroy")
            self.handle = ctypes.c_void_p(None)


class DeviceSlice:
    """(pointer, length) view of u64 field elements in device memory
    (crates/cudart/src/slice/device_slice.rs:11)."""

    def __init__(self, ptr, length):
        self.ptr = int(ptr)
        self.len = int(length)

    def __len__(self):
        return self.len

    def offset(self, elements):
        assert 0 <= elements <= self.len
        return DeviceSlice(self.ptr + 8 * elements, self.len - elements)

    def sub(self, start, length):
        assert 0 <= start and start + length <= self.len
        return DeviceSlice(self.ptr + 8 * start, length)

    def copy_from_host(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=np.uint64)
        assert arr.size <= self.len
        if stream is None:
            check(lib().ba_memcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D), "memcpy H2D")
        else:
            check(lib().ba_memcpy_async(self.ptr, arr.ctypes.data, arr.nbytes, H2D, stream.handle), "memcpy H2D")
            stream.synchronize()

    def to_host(self, count=None):
        count = self.len if count is None else count
        out = np.empty(count, dtype=np.uint64)
        check(lib().ba_memcpy(out.ctypes.data, self.ptr, 8 * count, D2H), "memcpy D2H")
        return out

    def copy_from_device(self, other, count=None):
        count = min(self.len, other.len) if count is None else count
        check(lib().ba_memcpy(self.ptr, other.ptr, 8 * count, D2D), "memcpy D2D")


class DeviceAllocation(DeviceSlice):
    """Owned device buffer of `length` u64 elements (crates/cudart/src/memory.rs:17)."""

    def __init__(self, length):
        p = ctypes.c_void_p(None)
        check(lib().ba_malloc(ctypes.byref(p), max(8, 8 * int(length))), "malloc")
        super().__init__(p.value, length)
        self._owned = True

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr, dtype=np.uint64)
        d = cls(arr.size)
        d.copy_from_host(arr)
        return d

    def free(self):
        if self._owned and self.ptr:
            check(lib().ba_free(self.ptr), "free")
            self.ptr = 0
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fill_uniform(dst, rows, cols, seed, stride=None, stream=None):
    """Seeded uniform field elements in HBM (include/boojum_amd.h: ba_fill_uniform)."""
    stride = rows if stride is None else stride
    h = stream.handle if stream is not None else None
    check(lib().ba_fill_uniform(dst.ptr, stride, rows, cols, seed, h), "fill_uniform")


def field_op(op, a, b, out, n, stream=None):
    h = stream.handle if stream is not None else None
    check(lib().ba_field_op(op, a.ptr, b.ptr if b is not None else None, out.ptr, n, h), "field_op")
