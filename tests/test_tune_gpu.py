"""GPU parity of tuned plans (ba_autotune / ba_tuning_get / ba_tuning_set): every pass split
and columns-per-workgroup choice the autotuner may pick gives results bit-identical to the
oracle, and invalid plans are rejected."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import log, raw_u64

LAT_MAX_LOG_ELEMS = 20  # batch NTTs of at most 2^20 elements run the latency-shaped passes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    for log_n in range(1, 29):
        for br in (False, True):
            boojum_amd.tune.set_plan(ctx, log_n, br, None)
    ctx.destroy()


def run_ntt(ba, x, log_n, cols, bitrev, inverse, le=0, ci=0):
    n = 1 << log_n
    d_in = ba.DeviceAllocation.from_host(x)
    d_out = ba.DeviceAllocation(cols * n)
    try:
        ba.ntt.batch_ntt_ctx(ba._test_ctx, d_in.ptr, d_out.ptr, log_n, cols, n, n, bitrev, inverse, le, ci)
        ba.synchronize()
        return d_out.to_host()
    finally:
        d_in.free()
        d_out.free()


def test_default_plans(ba):
    t = ba.tune
    assert t.get_plan(ba._test_ctx, 20) == ([12, 8], [0, 0])
    assert t.get_plan(ba._test_ctx, 13, True) == ([9, 4], [0, 0])
    assert t.get_plan(ba._test_ctx, 5) == ([5], [0])
    assert t.get_plan(ba._test_ctx, 28) == ([12, 8, 8], [0, 0, 0])


@pytest.mark.parametrize("log_n", [12, 13, 14, 16, 20, 21])
def test_every_split_bit_exact(ba, log_n):
    """Each split the autotuner can choose, with reps cycling through 0/1/2/4/8 and the
    contiguous pass's waves per SIMD through built-in/3/4, in both input orders, forward with a
    coset and inverse, against the oracle.  Batches are larger than 2^20 elements: smaller ones
    run the latency-shaped passes, which no plan selects."""
    cols = max(3, (1 << (LAT_MAX_LOG_ELEMS + 1) >> log_n) + 1)
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    x = raw_u64(rng, cols * n)
    splits = ba.tune.candidate_splits(log_n)
    assert splits, log_n
    modes = [(False, False, 1, 1), (True, True, 0, 0), (False, True, 2, 3), (True, False, 0, 0)]
    wants = {m: gl_oracle.batch_ntt(x, log_n, cols, n, n, m[0], m[1], m[2], m[3], nthreads=8) for m in modes}
    reps_cycle = [0, 1, 2, 4, 8]
    for k, bits in enumerate(splits):
        reps = [reps_cycle[(k + i) % 5] for i in range(len(bits))]
        for waves in ([0, 3, 4] if k < 2 else [(0, 3, 4)[k % 3]]):
            log(f"split log_n={log_n} bits={bits} reps={reps} contig_waves={waves}")
            for m in modes:
                ba.tune.set_plan(ba._test_ctx, log_n, m[0], bits, reps, contig_waves=waves)
                assert ba.tune.get_plan(ba._test_ctx, log_n, m[0]) == (bits, reps)
                assert ba.tune.get_contig_waves(ba._test_ctx, log_n, m[0]) == waves
                got = run_ntt(ba, x, log_n, cols, *m)
                assert np.array_equal(got, wants[m]), (bits, reps, waves, m)
    for br in (False, True):
        ba.tune.set_plan(ba._test_ctx, log_n, br, None)


@pytest.mark.parametrize("log_n,bits", [(19, [10, 9]), (19, [9, 10]), (22, [12, 10]), (21, [11, 10]),
                                         (22, [12, 5, 5])])
def test_wide_strided_passes_bit_exact(ba, log_n, bits):
    """The 9- and 10-bit strided passes (runs of 8 and 4 elements per row), alone above the
    contiguous pass or under another, in every mode including the LDE's gathered coset pass,
    against the oracle on sampled columns."""
    cols = 2
    n = 1 << log_n
    rng = np.random.default_rng(log_n + bits[-1])
    x = raw_u64(rng, cols * n)
    modes = [(False, False, 1, 1), (True, True, 0, 0), (False, True, 2, 3), (True, False, 3, 5)]
    for m in modes:
        log(f"wide strided log_n={log_n} bits={bits} mode={m}")
        ba.tune.set_plan(ba._test_ctx, log_n, m[0], bits, [0] * len(bits))
        got = run_ntt(ba, x, log_n, cols, *m)
        want = gl_oracle.batch_ntt(x, log_n, cols, n, n, m[0], m[1], m[2], m[3], nthreads=8)
        assert np.array_equal(got, want), (bits, m)
    # the LDE's coset NTTs gather the bit-reversed monomials in their first (top strided) pass
    ev = raw_u64(rng, n)
    d_in = ba.DeviceAllocation.from_host(ev)
    d_out = ba.DeviceAllocation(2 * n)
    try:
        ba.lde.lde_batch(ba._test_ctx, d_in, log_n, 1, 1, d_out)
        ba.synchronize()
        got = d_out.to_host()
        want = gl_oracle.lde(ev, log_n, 1, 1, nthreads=8)
        assert np.array_equal(got, want), bits
    finally:
        d_in.free()
        d_out.free()
    for br in (False, True):
        ba.tune.set_plan(ba._test_ctx, log_n, br, None)


@pytest.mark.parametrize("log_n", [13, 16])
def test_scaled_contiguous_pass_columns_per_workgroup(ba, log_n):
    """The two modes whose per-element coset factors are fused into the contiguous pass
    (forward b2n: before its first round; inverse n2b: after its last round) with 1, 2, 4 and 8
    columns per workgroup: each workgroup forms the factors anew for every column it repeats
    over, and every column matches the oracle."""
    cols = max(8, (1 << (LAT_MAX_LOG_ELEMS + 1) >> log_n) + 1)
    n = 1 << log_n
    rng = np.random.default_rng(100 + log_n)
    x = raw_u64(rng, cols * n)
    for bitrev, inverse in ((True, False), (False, True)):
        want = gl_oracle.batch_ntt(x, log_n, cols, n, n, bitrev, inverse, 2, 1, nthreads=8)
        bits, _ = ba.tune.get_plan(ba._test_ctx, log_n, bitrev)
        for r in (1, 2, 4, 8):
            log(f"scaled contiguous pass log_n={log_n} bitrev={bitrev} inverse={inverse} reps={r}")
            ba.tune.set_plan(ba._test_ctx, log_n, bitrev, bits, [r] * len(bits))
            got = run_ntt(ba, x, log_n, cols, bitrev, inverse, 2, 1)
            assert np.array_equal(got, want), (log_n, bitrev, inverse, r)
        ba.tune.set_plan(ba._test_ctx, log_n, bitrev, None)


def test_set_plan_rejects_invalid(ba):
    t = ba.tune
    bad = [
        (20, [12, 7]),        # does not cover 20 bits
        (20, [8, 8, 4]),      # contiguous pass below 9 bits for log_n > 12
        (20, [9, 11]),        # strided pass above 10 bits
        (20, [12, 3, 5]),     # strided pass below 4 bits
        (10, [6, 4]),         # log_n <= 12 is one pass
        (13, [13]),           # contiguous pass above 12 bits
    ]
    for log_n, bits in bad:
        with pytest.raises(ba.BoojumAmdError):
            t.set_plan(ba._test_ctx, log_n, False, bits)
    with pytest.raises(ba.BoojumAmdError):
        t.set_plan(ba._test_ctx, 20, False, [12, 8], [9, 0])  # reps above 8
    for waves in (1, 2, 5, 8):  # the contiguous pass runs at 3 or 4 waves per SIMD
        with pytest.raises(ba.BoojumAmdError):
            t.set_plan(ba._test_ctx, 20, False, [12, 8], [0, 0], contig_waves=waves)
    for log_n in (0, 29):
        with pytest.raises(ba.BoojumAmdError):
            t.get_plan(ba._test_ctx, log_n)
    # the LDE's coset NTTs: no strided pass above 8 bits (no gathered variant), no latency tiles
    with pytest.raises(ba.BoojumAmdError):
        t.set_plan(ba._test_ctx, 21, t.LDE_COSETS, [12, 9])
    with pytest.raises(ba.BoojumAmdError):
        t.set_plan(ba._test_ctx, 14, t.LDE_COSETS, [10, 4], tile_log=10)
    with pytest.raises(ba.BoojumAmdError):
        t.get_plan(ba._test_ctx, 14, 3)
    # a rejected plan leaves the size class unchanged
    assert t.get_plan(ba._test_ctx, 20) == ([12, 8], [0, 0])
    assert t.get_plan(ba._test_ctx, 21, t.LDE_COSETS) == ([12, 5, 4], [0, 0, 0])


def test_autotune_then_bit_exact(ba):
    """Tune 10..17 on small trial batches, then NTTs and an LDE of those sizes still match."""
    ba.tune.autotune(ba._test_ctx, 10, 17, elems_per_trial=1 << 20)
    for log_n in range(10, 18):
        for br in (False, True):
            bits, reps = ba.tune.get_plan(ba._test_ctx, log_n, br)
            log(f"tuned log_n={log_n} bitrev={br}: bits={bits} reps={reps}")
            assert sum(bits) == log_n and bits in ba.tune.candidate_splits(log_n)
            assert all(r in (0, 1, 2, 4, 8) for r in reps)
            assert ba.tune.get_tile_log(ba._test_ctx, log_n, br) in (0, 10)
        bits, reps = ba.tune.get_plan(ba._test_ctx, log_n, ba.tune.LDE_COSETS)
        log(f"tuned log_n={log_n} LDE cosets: bits={bits} reps={reps}")
        assert bits in ba.tune.candidate_splits(log_n) and all(b <= 8 for b in bits[1:])
        assert ba.tune.get_tile_log(ba._test_ctx, log_n, ba.tune.LDE_COSETS) == 0
    rng = np.random.default_rng(3)
    for log_n in (10, 13, 17):
        n = 1 << log_n
        cols = (1 << (LAT_MAX_LOG_ELEMS + 1) >> log_n) + 1  # past the latency-shaped passes
        x = raw_u64(rng, cols * n)
        for m in [(False, False, 1, 0), (True, True, 0, 0)]:
            want = gl_oracle.batch_ntt(x, log_n, cols, n, n, m[0], m[1], m[2], m[3], nthreads=8)
            assert np.array_equal(run_ntt(ba, x, log_n, cols, *m), want), (log_n, m)
        d_ev = ba.DeviceAllocation.from_host(x)
        d_out = ba.DeviceAllocation(cols * 4 * n)
        try:
            ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, 2, d_out)
            ba.synchronize()
            assert np.array_equal(d_out.to_host(), gl_oracle.lde(x, log_n, cols, 2, nthreads=8)), log_n
        finally:
            d_ev.free()
            d_out.free()


def test_autotune_rejects_bad_range(ba):
    for lo, hi in [(0, 5), (9, 8), (20, 29)]:
        with pytest.raises(ba.BoojumAmdError):
            ba.tune.autotune(ba._test_ctx, lo, hi, elems_per_trial=1 << 16)


def _names(ba, fn):
    from boojum_amd.profile import LaunchTrace
    with LaunchTrace(16) as tr:
        fn()
    ba.synchronize()
    names = list(tr.names)
    tr.destroy()
    return names


@pytest.mark.parametrize("log_n", [10, 14, 17])
def test_tile_log_choice(ba, log_n):
    """tile_log 10 runs the latency-shaped passes (1024-element tiles) at a batch size where the
    built-in rule would not, tile_log 12 runs the 4096-element tiles at one where it would; both
    match the oracle in every mode, and tile_log 0 restores the batch-size rule."""
    n = 1 << log_n
    big = (1 << (LAT_MAX_LOG_ELEMS + 1) >> log_n) + 1
    rng = np.random.default_rng(200 + log_n)
    modes = [(False, False, 1, 1), (True, True, 0, 0), (False, True, 2, 3), (True, False, 3, 5)]
    for tile, cols, want_lat in ((10, big, True), (12, 2, False), (0, 2, True), (0, big, False)):
        x = raw_u64(rng, cols * n)
        for m in modes:
            bits, reps = ba.tune.get_plan(ba._test_ctx, log_n, m[0])
            ba.tune.set_plan(ba._test_ctx, log_n, m[0], bits, reps, tile_log=tile)
            assert ba.tune.get_tile_log(ba._test_ctx, log_n, m[0]) == tile
            log(f"tile_log={tile} log_n={log_n} cols={cols} mode={m}")
            d_in = ba.DeviceAllocation.from_host(x)
            d_out = ba.DeviceAllocation(cols * n)
            try:
                names = _names(ba, lambda: ba.ntt.batch_ntt_ctx(ba._test_ctx, d_in.ptr, d_out.ptr, log_n, cols, n, n,
                                                                m[0], m[1], m[2], m[3]))
                assert names and all(("ntt_pass_lat" in s) == want_lat for s in names), (tile, cols, names)
                got = d_out.to_host()
            finally:
                d_in.free()
                d_out.free()
            want = gl_oracle.batch_ntt(x, log_n, cols, n, n, m[0], m[1], m[2], m[3], nthreads=8)
            assert np.array_equal(got, want), (tile, cols, m)
    for br in (False, True):
        ba.tune.set_plan(ba._test_ctx, log_n, br, None)


def test_tile_log_rejects_invalid(ba):
    t = ba.tune
    for log_n, tile in ((20, 10), (3, 10), (14, 11), (14, 8)):  # no latency shape above 18 / below 4
        bits, reps = t.get_plan(ba._test_ctx, log_n)
        with pytest.raises(ba.BoojumAmdError):
            t.set_plan(ba._test_ctx, log_n, False, bits, reps, tile_log=tile)
    assert t.get_tile_log(ba._test_ctx, 20) == 0


def test_batch_plan_takes_tuned_tiles(ba):
    """Batch plans run a merged NTT transform in the latency-shaped tiles only where its size class
    is tuned to them (tile_log 10), through their grouped kernels; results match the oracle."""
    log_n, cols = 14, 6
    n = 1 << log_n
    rng = np.random.default_rng(77)
    x = raw_u64(rng, cols * n)
    d_in = ba.DeviceAllocation.from_host(x)
    d_out = ba.DeviceAllocation(cols * n)
    try:
        for tile, want_lat in ((10, True), (0, False)):
            for br in (False, True):
                bits, reps = ba.tune.get_plan(ba._test_ctx, log_n, br)
                ba.tune.set_plan(ba._test_ctx, log_n, br, bits, reps, tile_log=tile)
            # three jobs: forward n2b (coset), inverse b2n, forward n2b (coset) again, 2 columns each
            jobs = [ba.batch.ntt_job(d_in.sub(2 * k * n, 2 * n), d_out.sub(2 * k * n, 2 * n), log_n, 2, n, n,
                                     bitrev_inputs=(k == 1), inverse=(k == 1), log_extension_degree=0 if k == 1 else 2,
                                     coset_index=0 if k == 1 else 1) for k in range(3)]
            plan = ba.batch.BatchPlan(ba._test_ctx, jobs)
            try:
                names = _names(ba, lambda: plan.launch())
            finally:
                plan.destroy()
            log(f"batch plan tile_log={tile}: {names}")
            assert names and any("ntt_pass_lat" in s for s in names) == want_lat, names
            got = d_out.to_host()
            for k in range(3):
                sl = slice(2 * k * n, 2 * (k + 1) * n)
                le, ci = (0, 0) if k == 1 else (2, 1)
                want = gl_oracle.batch_ntt(x[sl], log_n, 2, n, n, k == 1, k == 1, le, ci, nthreads=8)
                assert np.array_equal(got[sl], want), (tile, k)
    finally:
        d_in.free()
        d_out.free()
        for br in (False, True):
            ba.tune.set_plan(ba._test_ctx, log_n, br, None)


@pytest.mark.parametrize("log_n,bits", [(14, [9, 5]), (14, [10, 4]), (17, [9, 8]), (17, [9, 4, 4]), (20, [12, 4, 4]),
                                         (20, [11, 5, 4])])
def test_lde_coset_plans_bit_exact(ba, log_n, bits):
    """The LDE's coset NTTs under their own plan entry (gathered first pass of 4..8 bits, coset
    scale), with reps cycling, at blowups 2 and 4 against the oracle; the plain NTT plan of the
    class is left alone.  With three coset passes under a two-pass inverse NTT (17: [9, 4, 4]
    under [12, 5]) the cosets still fork onto the auxiliary streams."""
    cols = 3
    n = 1 << log_n
    x = raw_u64(np.random.default_rng(300 + log_n + bits[0]), cols * n)
    plain = ba.tune.get_plan(ba._test_ctx, log_n, False)
    try:
        for k, lb in enumerate((1, 2)):
            reps = [(1, 2, 4, 8)[(k + i) % 4] for i in range(len(bits))]
            ba.tune.set_plan(ba._test_ctx, log_n, ba.tune.LDE_COSETS, bits, reps)
            assert ba.tune.get_plan(ba._test_ctx, log_n, ba.tune.LDE_COSETS) == (bits, reps)
            assert ba.tune.get_plan(ba._test_ctx, log_n, False) == plain
            log(f"LDE coset plan log_n={log_n} bits={bits} reps={reps} blowup={1 << lb}")
            d_ev = ba.DeviceAllocation.from_host(x)
            d_out = ba.DeviceAllocation(cols * (n << lb))
            try:
                ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out)
                ba.synchronize()
                got = d_out.to_host()
            finally:
                d_ev.free()
                d_out.free()
            assert np.array_equal(got, gl_oracle.lde(x, log_n, cols, lb, nthreads=8)), (bits, lb)
    finally:
        ba.tune.set_plan(ba._test_ctx, log_n, ba.tune.LDE_COSETS, None)
