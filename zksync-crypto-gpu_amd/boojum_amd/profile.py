"""Per-launch kernel timing through the library's launch tracing (ba_profile_*)."""
import ctypes

from ._lib import check, lib
from .runtime import CudaEvent


class LaunchTrace:
    """Context manager: brackets every library kernel launch of this host thread with events."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.events = [CudaEvent() for _ in range(2 * capacity)]
        self._arr = (ctypes.c_void_p * (2 * capacity))(*[e.handle.value for e in self.events])
        self.count = 0

    def __enter__(self):
        check(lib().ba_profile_begin(self._arr, self.capacity), "profile_begin")
        return self

    def __exit__(self, *exc):
        n = ctypes.c_uint32(0)
        check(lib().ba_profile_end(ctypes.byref(n)), "profile_end")
        self.count = n.value
        self.names = [lib().ba_profile_launch_name(i).decode() for i in range(self.count)]

    def durations_ms(self):
        """[(kernel name, ms)] for the recorded launches (call after synchronising)."""
        return [(self.names[i], self.events[2 * i].elapsed_ms(self.events[2 * i + 1]))
                for i in range(self.count)]

    def destroy(self):
        for e in self.events:
            e.destroy()
