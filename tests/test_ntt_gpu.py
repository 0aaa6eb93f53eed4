"""GPU parity of the batch NTT against the CPU oracle, through the C-ABI.

Mirrors the reference's own test (crates/boojum-cuda/src/ntt.rs:391-663): for every log_n in
range, four device runs -- natural->bitrev in place and out of place, bitrev->natural in place
and out of place -- on columns at offset n inside a 2*n_max stride, compared element by element
with the CPU result.  The reference's #[ignore]d large cases are covered at the sizes the oracle
finishes in seconds, and to log_n = 24 through size-independent properties.
"""
import numpy as np
import pytest

import gl_oracle
from gpu_util import P, bitrev_perm, log, raw_u64

pytestmark = pytest.mark.gpu

NTTS_PER_BLOCK = 8  # the reference's batch granularity, used for its num_ntts choices


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    yield boojum_amd
    ctx.destroy()


def correctness(ba, log_n_range, inverse, log_extension_degree, coset_index, num_ntts, seed=0):
    n_max = 1 << (log_n_range[-1])
    OFFSET_MULTIPLIER = 1
    io_stride = n_max * (OFFSET_MULTIPLIER + 1)
    io_size = io_stride * num_ntts
    rng = np.random.default_rng(seed + 1000 * log_n_range[0] + 10 * int(inverse) + coset_index)
    inputs_host = raw_u64(rng, io_size)
    inputs_device = ba.DeviceAllocation(io_size)
    outputs_device = ba.DeviceAllocation(io_size)
    stream = ba.CudaStream()
    try:
        for log_n in log_n_range:
            log(f"ntt parity log_n={log_n} inverse={inverse} log_ext={log_extension_degree} "
                f"coset={coset_index} num_ntts={num_ntts}")
            n = 1 << log_n
            io_offset = n * OFFSET_MULTIPLIER
            perm = bitrev_perm(log_n)
            inputs_bitrev_host = inputs_host.copy()
            for k in range(io_size // n):
                chunk = inputs_host[k * n:(k + 1) * n]
                inputs_bitrev_host[k * n:(k + 1) * n][perm] = chunk

            results = {}
            for name, src, bitrev in (("n2b", inputs_host, False), ("b2n", inputs_bitrev_host, True)):
                inputs_device.copy_from_host(src)
                ba.ntt.batch_ntt_in_place(inputs_device, log_n, num_ntts, io_offset, io_stride, bitrev,
                                          inverse, log_extension_degree, coset_index, stream)
                stream.synchronize()
                results[name + "_in_place"] = inputs_device.to_host()
                inputs_device.copy_from_host(src)
                ba.ntt.batch_ntt_out_of_place(inputs_device, outputs_device, log_n, num_ntts, io_offset,
                                              io_offset, io_stride, io_stride, bitrev, inverse,
                                              log_extension_degree, coset_index, stream)
                stream.synchronize()
                results[name + "_out_of_place"] = outputs_device.to_host()

            # CPU oracle on the natural-order inputs: forward n2b / inverse natural->natural
            ref_n2b = gl_oracle.batch_ntt(inputs_host[io_offset:], log_n, num_ntts, io_stride, io_stride,
                                          False, inverse, log_extension_degree, coset_index, nthreads=8)
            ref_b2n = gl_oracle.batch_ntt(inputs_bitrev_host[io_offset:], log_n, num_ntts, io_stride,
                                          io_stride, True, inverse, log_extension_degree, coset_index,
                                          nthreads=8)
            for k in range(num_ntts):
                lo = k * io_stride
                for name, ref in (("n2b", ref_n2b), ("b2n", ref_b2n)):
                    want = ref[lo:lo + n]
                    for mode in ("_in_place", "_out_of_place"):
                        got = results[name + mode][lo + io_offset:lo + io_offset + n]
                        bad = np.nonzero(got != want)[0]
                        assert bad.size == 0, (f"{name}{mode} wrong for 2^{log_n}, ntt {k}, coset_index "
                                               f"{coset_index}, first index {bad[:4]}")
                # memory outside [offset, offset + n) of each column is untouched in place
                untouched = results["n2b_in_place"][lo:lo + io_offset]
                assert np.array_equal(untouched, inputs_host[lo:lo + io_offset])
    finally:
        inputs_device.free()
        outputs_device.free()
        stream.destroy()


def test_correctness_batch_ntt_fwd(ba):
    correctness(ba, list(range(1, 17)), False, 0, 0, 2 * NTTS_PER_BLOCK + 3)


def test_correctness_batch_lde_fwd(ba):
    correctness(ba, list(range(1, 17)), False, 3, 3, 3)


def test_correctness_batch_ntt_inv(ba):
    correctness(ba, list(range(1, 17)), True, 0, 0, 1)


def test_correctness_batch_lde_inv(ba):
    correctness(ba, list(range(1, 17)), True, 3, 3, 3)


def test_correctness_lde_intt_semantics(ba):
    # shivini's lde_intt: bit-reversed input, inverse, lde degree 2, coset 0 (primitives/ntt.rs:114-129)
    correctness(ba, [5, 12, 13, 17], True, 1, 0, 2)


def test_correctness_batch_ntt_fwd_large(ba):
    correctness(ba, list(range(17, 23)), False, 0, 0, 3)


def test_correctness_batch_lde_fwd_large(ba):
    correctness(ba, list(range(17, 22)), False, 3, 3, 3)


def test_correctness_batch_ntt_inv_large(ba):
    correctness(ba, list(range(17, 23)), True, 0, 0, 1)


def test_correctness_batch_lde_inv_large(ba):
    correctness(ba, list(range(17, 22)), True, 3, 3, 3)


def test_correctness_log_n_23_24(ba):
    correctness(ba, [23, 24], False, 1, 1, 1)


def test_golden_fixtures(ba, golden):
    n_max = 1 << 9
    d = ba.DeviceAllocation(n_max)
    o = ba.DeviceAllocation(n_max)
    try:
        for log_n in range(1, 10):
            n = 1 << log_n
            x = golden[f"in_{log_n}"]
            for b in (0, 1):
                for i in (0, 1):
                    for log_ext, coset_idx in [(0, 0), (3, 3), (1, 0)]:
                        d.copy_from_host(x)
                        ba.ntt.batch_ntt_internal(d.ptr, o.ptr, log_n, 1, n, n, b, i, log_ext, coset_idx)
                        ba.synchronize()
                        got = o.to_host(n)
                        assert np.array_equal(got, golden[f"ntt_{log_n}_{b}{i}_{log_ext}_{coset_idx}"]), \
                            (log_n, b, i, log_ext, coset_idx)
    finally:
        d.free()
        o.free()


@pytest.mark.parametrize("log_n", [20, 24, 26, 28])
def test_roundtrip_identity_full_size(ba, log_n):
    """forward (coset) then inverse (coset) is the identity; bit-exact at full size, up to the
    largest supported column (BA_MAX_LOG_N = 28, 2 GiB)."""
    n = 1 << log_n
    cols = 2 if log_n <= 24 else 1
    log(f"roundtrip log_n={log_n}")
    a = ba.DeviceAllocation(n * cols)
    b = ba.DeviceAllocation(n * cols)
    try:
        ba.runtime.fill_uniform(a, n, cols, 99)
        ba.ntt.batch_ntt_internal(a.ptr, b.ptr, log_n, cols, n, n, False, False, 2, 3)
        ba.ntt.batch_ntt_internal(b.ptr, b.ptr, log_n, cols, n, n, True, True, 2, 3)
        ba.synchronize()
        assert np.array_equal(a.to_host(), b.to_host())
        # b2n forward then n2b inverse also round-trips
        ba.ntt.batch_ntt_internal(a.ptr, b.ptr, log_n, cols, n, n, True, False, 0, 0)
        ba.ntt.batch_ntt_internal(b.ptr, b.ptr, log_n, cols, n, n, False, True, 0, 0)
        ba.synchronize()
        assert np.array_equal(a.to_host(), b.to_host())
    finally:
        a.free()
        b.free()


@pytest.mark.parametrize("log_n", [22, 28])
def test_delta_and_linearity_full_size(ba, log_n):
    n = 1 << log_n
    a = ba.DeviceAllocation(n)
    try:
        x = np.zeros(n, dtype=np.uint64)
        x[0] = 1
        a.copy_from_host(x)
        ba.ntt.batch_ntt_internal(a.ptr, a.ptr, log_n, 1, n, n, False, False, 0, 0)
        ba.synchronize()
        assert np.all(a.to_host() == 1)
        # delta at index 1 gives omega^bitrev(pos); spot-check against pow
        x[0], x[1] = 0, 1
        a.copy_from_host(x)
        ba.ntt.batch_ntt_internal(a.ptr, a.ptr, log_n, 1, n, n, False, False, 0, 0)
        ba.synchronize()
        y = a.to_host()
        w = gl_oracle.omega(log_n)
        for pos in np.random.default_rng(3).integers(0, n, 64):
            rev = int(format(int(pos), f"0{log_n}b")[::-1], 2)
            assert int(y[pos]) == pow(w, rev, P)
    finally:
        a.free()


def test_invalid_arguments_return_error(ba):
    a = ba.DeviceAllocation(64)
    try:
        cases = [
            (0, 1, 64, 64, 0, 0),    # log_n < 1
            (7, 1, 64, 64, 0, 0),    # n > stride
            (4, 1, 64, 64, 0, 1),    # coset without extension
            (4, 1, 64, 64, 2, 4),    # coset index out of range
            (28, 1, 2**31, 2**31, 3, 0),  # log_n + log_ext > 30
        ]
        for log_n, num, si, so, le, ci in cases:
            with pytest.raises(ba.BoojumAmdError) as e:
                ba.ntt.batch_ntt_internal(a.ptr, a.ptr, log_n, num, si, so, False, False, le, ci)
            assert e.value.code == 1
        with pytest.raises(ba.BoojumAmdError):
            ba.ntt.batch_ntt_out_of_place(a, a, 4, 1, 0, 0, 16, 16, False, False, 0, 0)  # aliasing
        with pytest.raises(ba.BoojumAmdError):  # overlapping (not identical) slices
            ba.ntt.batch_ntt_out_of_place(a.sub(0, 32), a.sub(16, 32), 4, 1, 0, 0, 16, 16, False, False, 0, 0)
        with pytest.raises(ba.BoojumAmdError):
            ba.ntt.batch_ntt_in_place(a, 4, 5, 0, 16, False, False, 0, 0)  # slice too short
        with pytest.raises(ba.BoojumAmdError):
            ba.ntt.batch_ntt_in_place(a, 4, 1, 1, 16, False, False, 0, 0)  # offset + n > stride
    finally:
        a.free()


def test_two_live_contexts(ba):
    """Explicit-context calls agree with the current-context entry point."""
    ctx2 = ba.Context.create(10, 14)
    n, cols, log_n = 1 << 14, 3, 14
    a = ba.DeviceAllocation(n * cols)
    b = ba.DeviceAllocation(n * cols)
    c = ba.DeviceAllocation(n * cols)
    try:
        ba.runtime.fill_uniform(a, n, cols, 5)
        ba.ntt.batch_ntt_internal(a.ptr, b.ptr, log_n, cols, n, n, False, False, 2, 1)
        ba.ntt.batch_ntt_ctx(ctx2, a.ptr, c.ptr, log_n, cols, n, n, False, False, 2, 1)
        ba.synchronize()
        assert np.array_equal(b.to_host(), c.to_host())
    finally:
        ctx2.destroy()
        for t in (a, b, c):
            t.free()


# Batches of at most 2^20 elements (log_n 4..18) run the latency-shaped passes (4 elements per
# thread, 1024-element tiles); larger batches the throughput passes.  The reference has no such
# split (ntt.rs:201-241 picks kernels by log_n alone); both shapes must give its results.
LAT_MAX_LOG_ELEMS = 20


def test_latency_passes_coset_single_column(ba):
    correctness(ba, list(range(4, 19)), False, 2, 1, 1, seed=7)


def test_latency_passes_inverse_coset_two_columns(ba):
    correctness(ba, list(range(4, 18)), True, 2, 3, 2, seed=8)


def _kernel_names(ba, fn):
    from boojum_amd.profile import LaunchTrace
    with LaunchTrace(16) as tr:
        fn()
    ba.synchronize()
    names = list(tr.names)
    tr.destroy()
    return names


@pytest.mark.parametrize("log_n,cols", [(16, 1), (12, 7), (18, 1), (18, 4), (18, 5), (16, 17), (19, 1)])
def test_latency_path_choice_and_agreement(ba, log_n, cols):
    """The path is chosen by batch size; at the boundary both paths give identical outputs
    (one call over all columns against per-column calls, which run the latency shape)."""
    n = 1 << log_n
    a = ba.DeviceAllocation(n * cols)
    b = ba.DeviceAllocation(n * cols)
    c = ba.DeviceAllocation(n * cols)
    try:
        ba.runtime.fill_uniform(a, n, cols, 11)
        for bitrev, inverse, le, ci in ((False, False, 0, 0), (True, True, 1, 1), (False, True, 2, 2),
                                        (True, False, 3, 5)):
            log(f"latency path log_n={log_n} cols={cols} bitrev={bitrev} inverse={inverse}")
            names = _kernel_names(ba, lambda: ba.ntt.batch_ntt_internal(a.ptr, b.ptr, log_n, cols, n, n, bitrev,
                                                                       inverse, le, ci))
            lat = (cols << log_n) <= (1 << LAT_MAX_LOG_ELEMS) and 4 <= log_n <= 18
            assert all(("ntt_pass_lat" in s) == lat for s in names), names
            for k in range(cols):
                ba.ntt.batch_ntt_internal(a.ptr + 8 * k * n, c.ptr + 8 * k * n, log_n, 1, n, n, bitrev, inverse,
                                          le, ci)
            ba.synchronize()
            assert np.array_equal(b.to_host(), c.to_host())
        # one column against the oracle
        x = a.to_host(n)
        want = gl_oracle.batch_ntt(x, log_n, 1, n, n, True, False, 3, 5, nthreads=8)
        assert np.array_equal(b.to_host(n), want)
    finally:
        for t in (a, b, c):
            t.free()
