"""boojum_cuda::context::Context mirror (crates/boojum-cuda/src/context.rs:157-296)."""
import ctypes

from ._lib import check, lib

OMEGA_LOG_ORDER = 30


class Context:
    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def create(cls, powers_of_w_coarse_log_count=12, powers_of_g_coarse_log_count=12):
        h = ctypes.c_void_p(None)
        check(lib().ba_context_create(powers_of_w_coarse_log_count, powers_of_g_coarse_log_count,
                                      ctypes.byref(h)), "Context::create")
        return cls(h)

    def destroy(self):
        if self.handle is not None and self.handle.value:
            check(lib().ba_context_destroy(self.handle), "Context::destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
