"""GPU parity of the benchmark's exact uneven batch: the 512 seeded jobs of bench.uneven_jobs
(log_n uniform in 12..22, random direction, 30% coset LDEs at blowup 2 or 4) planned on a context
autotuned for 12..22, as the benchmark runs them.  Every job is checked by a round trip through
the library's per-call path (the inverse transform gives back the input; an LDE's monomials
transform back to its evaluations) and every 16th job against the oracle; then the tuned plans of
the 18..22 size classes are checked against the oracle on their own."""
import os
import sys

import numpy as np
import pytest

import gl_oracle
from gpu_util import P, log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def tuned():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd.tune.autotune(ctx, 12, 22, elems_per_trial=1 << 26)
    yield boojum_amd, ctx
    ctx.destroy()


def test_bench_uneven_batch_on_tuned_context(tuned):
    ba, ctx = tuned
    sys.path.insert(0, ROOT)
    import bench
    jobs, calls, bufs, _ = bench.uneven_jobs(ba)
    plan = ba.batch.BatchPlan(ctx, jobs)
    try:
        plan.launch()
        ba.synchronize()
        checked = 0
        for i, (kind, d_in, d_out, d_mono, log_n, x) in enumerate(calls):
            n = 1 << log_n
            want_in = gl_oracle.fill_uniform(n, 1000 + i)
            tmp = ba.DeviceAllocation(n)
            try:
                if kind == "ntt":
                    # natural -> bit-reversed transform undone by the bit-reversed -> natural one
                    ba.ntt.batch_ntt_ctx(ctx, d_out.ptr, tmp.ptr, log_n, 1, n, n, True, not x)
                else:
                    # bit-reversed monomials -> natural evaluations (forward b2n)
                    ba.ntt.batch_ntt_ctx(ctx, d_mono.ptr, tmp.ptr, log_n, 1, n, n, True, False)
                ba.synchronize()
                assert np.array_equal(tmp.to_host(), want_in), ("round trip", i, kind, log_n)
            finally:
                tmp.free()
            if i % 16 == 0:
                got = d_out.to_host()
                if kind == "ntt":
                    want = gl_oracle.batch_ntt(want_in, log_n, 1, n, n, False, x, nthreads=4)
                else:
                    want = gl_oracle.lde(want_in, log_n, 1, x, nthreads=4)
                assert np.array_equal(got, want), ("oracle", i, kind, log_n)
                checked += 1
        log(f"uneven batch: {len(calls)} round trips, {checked} jobs against the oracle")
        assert checked >= 16
    finally:
        plan.destroy()
        for b in bufs:
            b.free()


@pytest.mark.parametrize("log_n", [18, 19, 20, 21, 22])
def test_tuned_plans_large_classes(tuned, log_n):
    """The plan the autotuner chose for each large size class, natural and bit-reversed input,
    forward and inverse with a coset, against the oracle (3 columns)."""
    ba, ctx = tuned
    n, cols = 1 << log_n, 3
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, 2**64, size=cols * n, dtype=np.uint64)
    for bitrev in (False, True):
        log(f"tuned plan 2^{log_n} bitrev={bitrev}: {ba.tune.get_plan(ctx, log_n, bitrev)}")
        for inverse in (False, True):
            d_in, d_out = ba.DeviceAllocation.from_host(x), ba.DeviceAllocation(cols * n)
            try:
                ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, cols, n, n, bitrev, inverse, 1, 1)
                ba.synchronize()
                got = d_out.to_host()
            finally:
                d_in.free()
                d_out.free()
            want = gl_oracle.batch_ntt(x, log_n, cols, n, n, bitrev, inverse, 1, 1, nthreads=4)
            assert np.array_equal(got, want), (log_n, bitrev, inverse)
