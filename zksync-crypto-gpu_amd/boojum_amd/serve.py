"""Serve mode: a request loop for coset-LDE jobs with memory-budget admission and continuous
batching (include/boojum_amd.h: ba_server_*).  No reference counterpart: shivini runs one
prover's LDEs in program order (crates/shivini/src/data_structures/storage.rs:154-260)."""
import ctypes

import numpy as np

from ._lib import ServeRequest, ServeResult, ServeStats, check, lib
from .runtime import DeviceSlice

INPUT_HOST, INPUT_DEVICE, INPUT_SYNTHETIC = 0, 1, 2


class Server:
    def __init__(self, ctx, memory_budget_bytes, max_inflight_batches=2):
        self.handle = ctypes.c_void_p(None)
        self._keep = {}  # host inputs must outlive their requests
        check(lib().ba_server_create(ctx.handle, memory_budget_bytes, max_inflight_batches,
                                     ctypes.byref(self.handle)), "server_create")

    def set_max_bypass(self, max_bypass):
        """Later requests that may be admitted ahead of a queued request waiting for memory
        (0: strict arrival order)."""
        check(lib().ba_server_set_max_bypass(self.handle, max_bypass), "server_set_max_bypass")

    def submit(self, log_n, num_cols, log_blowup, evals=None, seed=0, in_stride=0):
        """evals: a numpy uint64 array (copied in from host memory), a DeviceSlice (read in
        place), or None for seeded synthetic inputs.  Returns the ticket."""
        r = ServeRequest(log_n, num_cols, log_blowup, INPUT_SYNTHETIC, None, in_stride, seed)
        host = None
        if isinstance(evals, np.ndarray):
            host = np.ascontiguousarray(evals, dtype=np.uint64)
            r.input_kind, r.evals = INPUT_HOST, host.ctypes.data
        elif isinstance(evals, DeviceSlice):
            r.input_kind, r.evals = INPUT_DEVICE, evals.ptr
        elif evals is not None:
            raise TypeError("evals must be a numpy array, a DeviceSlice or None")
        t = ctypes.c_uint64(0)
        check(lib().ba_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)), "server_submit")
        if host is not None:
            self._keep[t.value] = host
        return t.value

    @staticmethod
    def _result(res, ticket):
        return {"ticket": ticket, "status": res.status, "outputs": res.outputs,
                "out_col_stride": res.out_col_stride, "queue_ms": res.queue_ms,
                "latency_ms": res.latency_ms, "batch": res.batch}

    def wait(self, ticket):
        res = ServeResult()
        check(lib().ba_server_wait(self.handle, ticket, ctypes.byref(res)), "server_wait")
        self._keep.pop(ticket, None)
        return self._result(res, ticket)

    def poll(self, ticket):
        done, res = ctypes.c_int(0), ServeResult()
        check(lib().ba_server_poll(self.handle, ticket, ctypes.byref(done), ctypes.byref(res)), "server_poll")
        if not done.value:
            return None
        self._keep.pop(ticket, None)
        return self._result(res, ticket)

    def output(self, result, log_n, num_cols, log_blowup):
        """The outputs of a finished request as a DeviceSlice (column-major, B * n per column)."""
        return DeviceSlice(result["outputs"], num_cols << (log_n + log_blowup))

    def release(self, ticket):
        check(lib().ba_server_release(self.handle, ticket), "server_release")

    def stats(self):
        st = ServeStats()
        check(lib().ba_server_stats(self.handle, ctypes.byref(st)), "server_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def destroy(self):
        if self.handle is not None and self.handle.value:
            check(lib().ba_server_destroy(self.handle), "server_destroy")
            self.handle = None
        self._keep.clear()
