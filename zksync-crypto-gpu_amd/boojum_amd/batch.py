"""Uneven batches of NTT and LDE jobs (include/boojum_amd.h: ba_batch_plan_*)."""
import ctypes

from ._lib import Job, check, lib

JOB_NTT, JOB_LDE = 0, 1


def ntt_job(inputs, outputs, log_n, num_cols, stride_in, stride_out, bitrev_inputs=False, inverse=False,
            log_extension_degree=0, coset_index=0):
    return Job(JOB_NTT, log_n, num_cols, int(bitrev_inputs), int(inverse), log_extension_degree,
               coset_index, 0, inputs.ptr, outputs.ptr, None, stride_in, stride_out, 0)


def lde_job(evals, outputs, log_n, num_cols, log_blowup, stride_in=None, stride_out=None,
            out_coset_stride=None, monomials=None):
    n = 1 << log_n
    return Job(JOB_LDE, log_n, num_cols, 0, 0, 0, 0, log_blowup, evals.ptr, outputs.ptr,
               monomials.ptr if monomials is not None else None, n if stride_in is None else stride_in,
               (n << log_blowup) if stride_out is None else stride_out,
               n if out_coset_stride is None else out_coset_stride)


class BatchPlan:
    def __init__(self, ctx, jobs):
        arr = (Job * len(jobs))(*jobs)
        self.handle = ctypes.c_void_p(None)
        check(lib().ba_batch_plan_create(ctx.handle, arr, len(jobs), ctypes.byref(self.handle)), "batch_plan_create")

    def launch(self, stream=None):
        check(lib().ba_batch_plan_launch(self.handle, stream.handle if stream is not None else None),
              "batch_plan_launch")

    def stats(self):
        l, v, w = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_uint64(0)
        check(lib().ba_batch_plan_stats(self.handle, ctypes.byref(l), ctypes.byref(v), ctypes.byref(w)), "stats")
        return {"launches": l.value, "levels": v.value, "workgroups": w.value}

    def destroy(self):
        if self.handle is not None and self.handle.value:
            check(lib().ba_batch_plan_destroy(self.handle), "batch_plan_destroy")
            self.handle = None
