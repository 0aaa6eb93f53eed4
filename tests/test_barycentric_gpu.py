"""GPU parity of barycentric evaluation (ba_barycentric_*), mirroring the reference's tests
(crates/boojum-cuda/src/barycentric.rs:300-683: coefficients at base and extension points for
log_count 0..19, batch evaluation of base/extension columns) against the oracle, plus the LDE
consistency check: every bit-reversed coset of an LDE evaluates to the same value at X."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import P, log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    ctx.destroy()


def rnd(rng):
    return int(rng.integers(0, 2**64, dtype=np.uint64))


def gpu_coeffs(ba, x, coset, count, bit_reverse, ext):
    d_x = ba.DeviceAllocation.from_host(np.array(list(x) if ext else [x], dtype=np.uint64))
    d_cf = ba.DeviceAllocation(2)
    d_l = ba.DeviceAllocation(count * (2 if ext else 1))
    try:
        ba.barycentric.precompute_lagrange_coeffs(d_x, d_cf, coset, d_l, bit_reverse, ext=ext)
        ba.synchronize()
        return d_l.to_host()
    finally:
        for t in (d_x, d_cf, d_l):
            t.free()


def test_precompute_lagrange_coeffs(ba):
    rng = np.random.default_rng(1)
    coset = 7
    for log_count in list(range(0, 12)) + [16, 19]:
        count = 1 << log_count
        for ext in (False, True):
            for bitrev in (False, True):
                x = (rnd(rng), rnd(rng)) if ext else rnd(rng)
                log(f"coeffs log_count={log_count} ext={ext} bitrev={bitrev}")
                got = gpu_coeffs(ba, x, coset, count, bitrev, ext)
                xo = (x[0] % P, x[1] % P) if ext else x % P
                assert np.array_equal(got, gl_oracle.barycentric_coeffs(xo, coset, count, bitrev, ext=ext)), \
                    (log_count, ext, bitrev)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_batch_eval(ba, kind):
    rng = np.random.default_rng(2 + kind)
    for log_count, num_polys in [(0, 3), (4, 1), (12, 5), (13, 9), (16, 20), (20, 4)]:
        count = 1 << log_count
        ext = kind != 0
        x = (rnd(rng), rnd(rng)) if ext else rnd(rng)
        xo = (x[0] % P, x[1] % P) if ext else x % P
        coeffs = gl_oracle.barycentric_coeffs(xo, 7, count, True, ext=ext)
        ys = raw_u64(rng, count * num_polys * (2 if kind == 2 else 1))
        want = gl_oracle.barycentric_eval(kind, ys, count, num_polys, coeffs)
        log(f"batch_eval kind={kind} log_count={log_count} polys={num_polys}")
        temp_n = ba.barycentric.batch_eval_temp_elems(kind, count, num_polys)
        d_ys, d_l = ba.DeviceAllocation.from_host(ys), ba.DeviceAllocation.from_host(coeffs)
        d_t, d_e = ba.DeviceAllocation(temp_n), ba.DeviceAllocation(len(want))
        try:
            ba.barycentric.batch_eval(kind, d_ys, count, num_polys, d_l, d_t, d_e)
            ba.synchronize()
            assert np.array_equal(d_e.to_host(), want), (kind, log_count, num_polys)
        finally:
            for t in (d_ys, d_l, d_t, d_e):
                t.free()


def test_lde_cosets_agree(ba):
    """GPU LDE of 6 columns of 2^14 at blowup 4, then GPU barycentric on the evaluations
    (natural, coset 1) and on each bit-reversed coset (shift 7 w_{Bn}^bitrev(c)): all agree."""
    log_n, lb, cols = 14, 2, 6
    n, B = 1 << log_n, 1 << lb
    rng = np.random.default_rng(5)
    ev = raw_u64(rng, n * cols)
    d_ev, d_out = ba.DeviceAllocation.from_host(ev), ba.DeviceAllocation(B * n * cols)
    d_mono = ba.DeviceAllocation(n * cols)
    # coset-major output: coset c of every column is one matrix, as shivini stores cosets
    ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono, out_col_stride=n,
                     out_coset_stride=cols * n)
    ba.synchronize()
    x = (rnd(rng), rnd(rng))
    temp_n = ba.barycentric.batch_eval_temp_elems(1, n, cols)
    d_t, d_e = ba.DeviceAllocation(temp_n), ba.DeviceAllocation(2 * cols)
    d_x = ba.DeviceAllocation.from_host(np.array(x, dtype=np.uint64))
    d_cf, d_l = ba.DeviceAllocation(2), ba.DeviceAllocation(2 * n)
    W30 = 0x7E9BD009B86A0845
    try:
        def evaluate(ys, coset, bitrev):
            ba.barycentric.precompute_lagrange_coeffs(d_x, d_cf, coset, d_l, bitrev, ext=True)
            ba.barycentric.batch_eval(1, ys, n, cols, d_l, d_t, d_e)
            ba.synchronize()
            return d_e.to_host()

        want = evaluate(d_ev, 1, False)
        xo = (x[0] % P, x[1] % P)
        assert np.array_equal(want, gl_oracle.barycentric_eval(1, ev, n, cols,
                                                               gl_oracle.barycentric_coeffs(xo, 1, n, False, ext=True)))
        for c in range(B):
            rev = int(format(c, f"0{lb}b")[::-1], 2)
            shift = 7 * pow(W30, rev << (30 - log_n - lb), P) % P
            got = evaluate(d_out.sub(c * cols * n, cols * n), shift, True)
            assert np.array_equal(got, want), c
    finally:
        for t in (d_ev, d_out, d_mono, d_t, d_e, d_x, d_cf, d_l):
            t.free()


def test_bad_arguments(ba):
    d = ba.DeviceAllocation(64)
    try:
        with pytest.raises(ba.BoojumAmdError):  # count not a power of two
            ba.check(ba.lib().ba_barycentric_precompute_lagrange_coeffs(0, d.ptr, d.ptr, 7, d.ptr, 12, 0, None))
        with pytest.raises(ba.BoojumAmdError):  # temp storage too small
            ba.check(ba.lib().ba_barycentric_batch_eval(0, d.ptr, 16, 2, d.ptr, d.ptr, 0, d.ptr, None))
        with pytest.raises(ba.BoojumAmdError):
            ba.barycentric.batch_eval_temp_elems(3, 16, 1)
    finally:
        d.free()
