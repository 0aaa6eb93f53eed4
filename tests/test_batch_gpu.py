"""GPU parity of uneven batch plans (ba_batch_plan_*): every job of a mixed batch -- NTTs of
different sizes, orders, directions and cosets, and coset LDEs -- matches the oracle run on that
job alone, plan after plan and replay after replay."""
import ctypes

import numpy as np
import pytest

import gl_oracle
from gpu_util import log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    ctx.destroy()


def make_batch(ba, rng, count, log_n_choices, lde_fraction=0.3):
    """Random jobs with their host inputs and device buffers."""
    jobs = []
    for i in range(count):
        log_n = int(rng.choice(log_n_choices))
        n = 1 << log_n
        cols = int(rng.integers(1, 4))
        if rng.random() < lde_fraction and log_n <= 24:
            lb = int(rng.integers(1, 3))
            stride = n + int(rng.integers(0, 3))
            x = raw_u64(rng, cols * stride)
            d_in = ba.DeviceAllocation.from_host(x)
            d_out = ba.DeviceAllocation(cols * (n << lb))
            d_mono = ba.DeviceAllocation(cols * n)
            job = ba.batch.lde_job(d_in, d_out, log_n, cols, lb, stride_in=stride, monomials=d_mono)
            jobs.append(dict(kind="lde", log_n=log_n, cols=cols, lb=lb, stride=stride, x=x, job=job,
                             bufs=(d_in, d_out, d_mono), out=d_out))
        else:
            bitrev, inverse = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
            le = int(rng.integers(0, 3))
            ci = int(rng.integers(0, 1 << le)) if le else 0
            stride = n + int(rng.integers(0, 3))
            x = raw_u64(rng, cols * stride)
            d_in = ba.DeviceAllocation.from_host(x)
            d_out = ba.DeviceAllocation(cols * stride)
            job = ba.batch.ntt_job(d_in, d_out, log_n, cols, stride, stride, bitrev, inverse, le, ci)
            jobs.append(dict(kind="ntt", log_n=log_n, cols=cols, stride=stride, x=x, job=job, bitrev=bitrev,
                             inverse=inverse, le=le, ci=ci, bufs=(d_in, d_out), out=d_out))
    return jobs


def check_jobs(jobs):
    for j in jobs:
        n = 1 << j["log_n"]
        got = j["out"].to_host()
        if j["kind"] == "lde":
            want = gl_oracle.lde(j["x"], j["log_n"], j["cols"], j["lb"], in_stride=j["stride"], nthreads=4)
            assert np.array_equal(got, want), ("lde", j["log_n"], j["cols"], j["lb"])
        else:
            want = gl_oracle.batch_ntt(j["x"], j["log_n"], j["cols"], j["stride"], j["stride"], j["bitrev"],
                                       j["inverse"], j["le"], j["ci"], nthreads=4)
            for c in range(j["cols"]):
                s = c * j["stride"]
                assert np.array_equal(got[s:s + n], want[s:s + n]), ("ntt", j["log_n"], j["bitrev"], j["inverse"])


def free(jobs):
    for j in jobs:
        for b in j["bufs"]:
            b.free()


@pytest.mark.parametrize("seed", [0, 1])
def test_mixed_batch_small(ba, seed):
    rng = np.random.default_rng(seed)
    jobs = make_batch(ba, rng, 48, list(range(1, 17)))
    log(f"mixed batch seed={seed}: {len(jobs)} jobs")
    plan = ba.batch.BatchPlan(ba._test_ctx, [j["job"] for j in jobs])
    try:
        st = plan.stats()
        assert st["launches"] < sum(1 for _ in jobs) * 3  # grouped, not one launch per pass
        plan.launch()
        ba.synchronize()
        check_jobs(jobs)
    finally:
        plan.destroy()
        free(jobs)


def test_mixed_batch_config_sizes(ba):
    """log_n 12..22 as in the uneven-batch configuration, fewer jobs."""
    rng = np.random.default_rng(12)
    jobs = make_batch(ba, rng, 24, list(range(12, 23)))
    log(f"mixed batch large: {[j['log_n'] for j in jobs]}")
    plan = ba.batch.BatchPlan(ba._test_ctx, [j["job"] for j in jobs])
    try:
        plan.launch()
        ba.synchronize()
        check_jobs(jobs)
    finally:
        plan.destroy()
        free(jobs)


def test_plan_graph_replay(ba):
    """A plan captured in a hipGraph gives the same results on every replay (inputs restored
    between replays by device copies; NTT jobs out of place, LDE jobs keep their evaluations)."""
    rng = np.random.default_rng(5)
    jobs = make_batch(ba, rng, 16, list(range(4, 15)), lde_fraction=0.4)
    plan = ba.batch.BatchPlan(ba._test_ctx, [j["job"] for j in jobs])
    lib = ba.lib()
    stream = ba.CudaStream()
    try:
        ba.check(lib.ba_stream_begin_capture(stream.handle), "capture")
        plan.launch(stream)
        graph = ctypes.c_void_p(None)
        ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(graph)), "end capture")
        for rep in range(3):
            for j in jobs:
                j["out"].copy_from_host(np.zeros(len(j["out"]), dtype=np.uint64))
            ba.check(lib.ba_graph_launch(graph, stream.handle), "graph launch")
            stream.synchronize()
            check_jobs(jobs)
        ba.check(lib.ba_graph_destroy(graph), "graph destroy")
    finally:
        plan.destroy()
        free(jobs)
        stream.destroy()


def test_plan_rejects_bad_job(ba):
    d = ba.DeviceAllocation(64)
    try:
        bad = ba.batch.ntt_job(d, d, 7, 1, 64, 64)  # n > stride
        with pytest.raises(ba.BoojumAmdError):
            ba.batch.BatchPlan(ba._test_ctx, [bad])
    finally:
        d.free()


def test_merged_columns(ba):
    """Jobs sharing a transform are merged into one multi-column launch per pass over columns in
    separate buffers: in-place and out-of-place NTTs, multi-column jobs with padded strides, and
    LDEs with and without a monomial buffer, all matching the oracle job by job."""
    rng = np.random.default_rng(77)
    jobs = []
    for i in range(36):
        log_n = (3, 10, 13, 16)[i % 4]
        n = 1 << log_n
        cols = 1 + (i % 3)
        stride = n + (i % 2)
        x = raw_u64(rng, cols * stride)
        d_in = ba.DeviceAllocation.from_host(x)
        if i % 6 < 4:
            in_place = i % 2 == 0
            d_out = d_in if in_place else ba.DeviceAllocation(cols * stride)
            job = ba.batch.ntt_job(d_in, d_out, log_n, cols, stride, stride, False, False, 1, 1)
            jobs.append(dict(kind="ntt", log_n=log_n, cols=cols, stride=stride, x=x, job=job, bitrev=False,
                             inverse=False, le=1, ci=1, bufs=(d_in,) if in_place else (d_in, d_out), out=d_out))
        else:
            d_out = ba.DeviceAllocation(cols * (n << 1))
            d_mono = ba.DeviceAllocation(cols * n) if i % 2 else None
            job = ba.batch.lde_job(d_in, d_out, log_n, cols, 1, stride_in=stride, monomials=d_mono)
            bufs = (d_in, d_out) + ((d_mono,) if d_mono is not None else ())
            jobs.append(dict(kind="lde", log_n=log_n, cols=cols, lb=1, stride=stride, x=x, job=job,
                             bufs=bufs, out=d_out))
    plan = ba.batch.BatchPlan(ba._test_ctx, [j["job"] for j in jobs])
    try:
        st = plan.stats()
        log(f"merged batch: {st}")
        # 4 sizes x (NTT, LDE) transforms: fewer grouped launches than jobs
        assert st["launches"] < len(jobs), st
        plan.launch()
        ba.synchronize()
        check_jobs(jobs)
    finally:
        plan.destroy()
        free(jobs)
