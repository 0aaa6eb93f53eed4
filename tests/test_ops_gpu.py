"""GPU parity of the bit-reverse permutation, the device field arithmetic and the seeded input
generator, through the C-ABI."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import P, bitrev_perm, column_seed, log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    return boojum_amd


@pytest.mark.parametrize("in_place", [False, True])
@pytest.mark.parametrize("log_rows", [0, 1, 5, 9, 10, 11, 16, 21])
def test_bit_reverse(ba, in_place, log_rows):
    # crates/boojum-cuda/src/ops_complex.rs:1628-1672 uses LOG_ROWS = 16, COLS = 16
    log(f"bit_reverse log_rows={log_rows} in_place={in_place}")
    rows = 1 << log_rows
    cols = 16 if log_rows <= 16 else 2
    stride = rows + 3  # strided columns, as a DeviceMatrix chunk may be
    rng = np.random.default_rng(log_rows)
    h_src = raw_u64(rng, cols * stride)
    d_src = ba.DeviceAllocation.from_host(h_src)
    d_dst = ba.DeviceAllocation(cols * stride)
    try:
        if in_place:
            ba.ops.bit_reverse_in_place(d_src, log_rows, cols, stride=stride)
            got = d_src.to_host()
        else:
            d_dst.copy_from_host(np.zeros(cols * stride, dtype=np.uint64))
            ba.ops.bit_reverse(d_src, d_dst, log_rows, cols, src_stride=stride, dst_stride=stride)
            got = d_dst.to_host()
        ba.synchronize()
        perm = bitrev_perm(log_rows)
        for c in range(cols):
            s = h_src[c * stride:c * stride + rows]
            d = got[c * stride:c * stride + rows]
            assert np.array_equal(d[perm], s)
            if in_place:  # padding between columns untouched
                assert np.array_equal(got[c * stride + rows:(c + 1) * stride], h_src[c * stride + rows:(c + 1) * stride])
    finally:
        d_src.free()
        d_dst.free()


def test_bit_reverse_rejects_overlap(ba):
    d = ba.DeviceAllocation(1 << 12)
    try:
        with pytest.raises(ba.BoojumAmdError):
            ba.ops.bit_reverse(d, d.offset(8), 10, 2)
    finally:
        d.free()


def test_field_ops(ba):
    rng = np.random.default_rng(7)
    edges = [0, 1, 2, P - 2, P - 1, P, P + 1, 2**64 - 1, 2**32 - 1, 2**32, 2**63, 2**64 - 2**32]
    a = np.array(edges * len(edges) + list(raw_u64(rng, 4096)), dtype=np.uint64)
    b = np.array([e for e in edges for _ in edges] + list(raw_u64(rng, 4096)), dtype=np.uint64)
    n = a.size
    da, db, do = ba.DeviceAllocation.from_host(a), ba.DeviceAllocation.from_host(b), ba.DeviceAllocation(n)
    try:
        for op, fn in [(0, lambda x, y: (x + y) % P), (1, lambda x, y: (x - y) % P),
                       (2, lambda x, y: (x * y) % P), (3, lambda x, y: x % P)]:
            ba.runtime.field_op(op, da, db, do, n)
            got = do.to_host()
            for i in range(n):
                assert int(got[i]) == fn(int(a[i]), int(b[i])), (op, int(a[i]), int(b[i]))
        ba.runtime.field_op(4, da, None, do, 256)
        got = do.to_host(256)
        for i in range(256):
            x = int(a[i]) % P
            assert int(got[i]) == (pow(x, P - 2, P) if x else 0)
    finally:
        for t in (da, db, do):
            t.free()


def _run_op(ba, op, a, b):
    da, db, do = ba.DeviceAllocation.from_host(a), ba.DeviceAllocation.from_host(b), ba.DeviceAllocation(a.size)
    try:
        ba.runtime.field_op(op, da, db, do, a.size)
        return do.to_host()
    finally:
        for t in (da, db, do):
            t.free()


SHIFT_EDGES = [0, 1, 2, P - 2, P - 1, P, P + 1, 2**64 - 1, 2**32 - 1, 2**32, 2**63, 2**64 - 2**32,
               2**63 - 1, 2**48, 2**16 - 1, 2**48 - 1, 2**64 - 2**48, 0xFFFFFFFF00000000,
               0x0000FFFFFFFFFFFF, 0xFFFF000000000000, 0x00000000FFFF0000, 0x8000000080000000]


def test_field_shifts(ba):
    """x 2^E for every shift the NTT blocks use (multiples of 12 below 192) and the boundaries of
    the three shift forms, on edge values of x (any u64, canonical or not) and random ones."""
    exps = [12, 24, 36, 48, 60, 72, 84, 96, 108, 120, 132, 144, 156, 168, 180, 0, 1, 31, 32, 33, 63, 64,
            65, 191]
    rng = np.random.default_rng(11)
    xs = [x for x in SHIFT_EDGES for _ in range(24)] + list(raw_u64(rng, 4096))
    # small multiples of 2^(64-F) and 2^(96-F) make the shifted words wrap or vanish exactly
    for f in (12, 24, 36, 48, 60):
        xs += [k << (64 - f) for k in (1, 2, 3, 2**f - 1)] + [(k << (96 - f)) % 2**64 for k in (1, 5)]
    sel = [i % 24 for i in range(24 * len(SHIFT_EDGES))] + list(rng.integers(0, 24, len(xs) - 24 * len(SHIFT_EDGES)))
    # a factor 2^0 or 2^96 = -1 is no shift: those take canonical x, as everywhere they are used
    xs = [x % P if exps[k] % 96 == 0 else x for x, k in zip(xs, sel)]
    a = np.array(xs, dtype=np.uint64)
    b = np.array(sel, dtype=np.uint64)
    got = _run_op(ba, 5, a, b)
    for i in range(a.size):
        e = exps[int(b[i]) % 24]
        assert int(got[i]) == (int(a[i]) << e) % P, ("pow2", hex(int(a[i])), e)


def test_field_reductions(ba):
    """The word reductions on chosen inputs that reach every carry/borrow case: lo + hl (2^32 - 1)
    wrapping or not, the subtraction of hh (+ a borrow-in) going below zero or not, and results at
    or above p, against (lo + hl 2^64 + (hh + bi) 2^96) mod p."""
    rng = np.random.default_rng(13)
    E = 2**32 - 1
    los, words = [], []
    for _ in range(64):
        hl = int(rng.integers(0, 2**32))
        hh = int(rng.integers(1, 2**31))
        r = int(rng.integers(0, hh))
        wrap = (2**64 - hl * E) % 2**64
        for lo in (r, (wrap + r) % 2**64, (wrap + hh + r) % 2**64, (wrap + 2**64 - 1) % 2**64, 2**64 - 1 - r,
                   P + r, int(rng.integers(0, 2**64, dtype=np.uint64))):
            for hh_ in (hh, 0, 1):
                for bi in (0, 1):
                    los.append(lo)
                    words.append(hl | (hh_ << 32) | (bi << 63))
    # hl = 0: R = lo exactly; hh above it borrows
    for lo in (0, 1, 5, P - 1, P, 2**64 - 1):
        for hh_ in (0, 1, 7, 2**31 - 1):
            for bi in (0, 1):
                los.append(lo)
                words.append((hh_ << 32) | (bi << 63))
    a = np.array(los, dtype=np.uint64)
    b = np.array(words, dtype=np.uint64)
    got = _run_op(ba, 6, a, b)
    for i in range(a.size):
        lo, w = int(a[i]), int(b[i])
        hl, hh, bi = w & 0xFFFFFFFF, (w >> 32) & 0x7FFFFFFF, w >> 63
        want = (lo + hl * 2**64 + (hh + bi) * 2**96) % P
        assert int(got[i]) == want, ("reduce3b", hex(lo), hl, hh, bi)
    # reduce3: full 32-bit hh, no borrow-in
    b3 = np.array([w & 0x7FFFFFFFFFFFFFFF | ((w & 1) << 63) for w in (int(x) for x in b)], dtype=np.uint64)
    got = _run_op(ba, 8, a, b3)
    for i in range(a.size):
        lo, w = int(a[i]), int(b3[i])
        hl, hh = w & 0xFFFFFFFF, w >> 32
        assert int(got[i]) == (lo + hl * 2**64 + hh * 2**96) % P, ("reduce3", hex(lo), hl, hh)
    got = _run_op(ba, 7, a, b)
    for i in range(a.size):
        lo, hl = int(a[i]), int(b[i]) & 0xFFFFFFFF
        assert int(got[i]) == (lo + hl * 2**64) % P, ("reduce2", hex(lo), hl)


def test_field_mul_edges(ba):
    """Products whose 128-bit words make the reduction borrow or wrap: powers of two times
    multiples of 2^48, 2^32 and 2^16 (x 2^96 = -x), and edge values against each other."""
    vals = SHIFT_EDGES + [(k << s) % 2**64 for s in (16, 32, 48, 56) for k in (1, 3, 2**16 - 1)]
    pairs = [(x, y) for x in vals for y in vals]
    a = np.array([x for x, _ in pairs], dtype=np.uint64)
    b = np.array([y for _, y in pairs], dtype=np.uint64)
    got = _run_op(ba, 2, a, b)
    for i in range(a.size):
        assert int(got[i]) == (int(a[i]) * int(b[i])) % P, ("mul", hex(int(a[i])), hex(int(b[i])))


def test_fill_uniform_matches_oracle(ba):
    rows, cols, seed = 1000, 3, 12345
    d = ba.DeviceAllocation(rows * cols)
    try:
        ba.runtime.fill_uniform(d, rows, cols, seed)
        got = d.to_host()
        for k in range(cols):
            assert np.array_equal(got[k * rows:(k + 1) * rows], gl_oracle.fill_uniform(rows, column_seed(seed, k)))
        assert np.all(got < np.uint64(P))
    finally:
        d.free()


# ---- powers and omega shifts (ops_complex.rs:31-260) against the oracle ----
@pytest.fixture(scope="module")
def tables(ba):
    """The power tables these calls read: the current device's live context."""
    ctx = ba.Context.create(12, 12)
    yield ctx
    ctx.destroy()

POWER_CASES = [  # (log_degree, offset, inverse, bit_reverse, count)
    (0, 0, False, False, 5), (0, 3, True, True, 5), (1, 0, False, True, 2), (10, 0, False, False, 1024),
    (10, 0, True, True, 1024), (16, 12345, False, True, 70000), (20, 2**32 - 1000, True, False, 3000),
    (30, 7, False, False, 4099), (30, 0, True, True, 1 << 16), (12, 2**31, False, True, 777),
]


def test_powers_of_w_and_g(ba, tables):
    for ld, off, inv, bitrev, count in POWER_CASES:
        log(f"powers of w/g log_degree={ld} offset={off} inverse={inv} bitrev={bitrev} count={count}")
        d = ba.DeviceAllocation(count)
        try:
            ba.ops.get_powers_of_w(ld, off, inv, bitrev, d)
            ba.synchronize()
            assert np.array_equal(d.to_host(), gl_oracle.powers_of_w(ld, off, inv, bitrev, count))
            ba.ops.get_powers_of_g(ld, off, inv, bitrev, d)
            ba.synchronize()
            assert np.array_equal(d.to_host(), gl_oracle.powers_of_g(ld, off, inv, bitrev, count))
        finally:
            d.free()


def test_powers_by_val_and_ref(ba):
    rng = np.random.default_rng(8)
    for off, bitrev, count in [(0, False, 1), (0, False, 100000), (5, True, 3000), (2**32 - 50, False, 4000),
                               (123, False, 1 << 20)]:
        base = int(rng.integers(0, 2**64, dtype=np.uint64))
        e0, e1 = (int(x) for x in rng.integers(0, 2**64, size=2, dtype=np.uint64))
        log(f"powers by val/ref offset={off} bitrev={bitrev} count={count}")
        d = ba.DeviceAllocation(2 * count)
        b = ba.DeviceAllocation.from_host(np.array([base, 0], dtype=np.uint64))
        be = ba.DeviceAllocation.from_host(np.array([e0, e1], dtype=np.uint64))
        try:
            want = gl_oracle.powers_by_val_bf(base, off, bitrev, count)
            ba.ops.get_powers_by_val_bf(base, off, bitrev, d.sub(0, count))
            ba.synchronize()
            assert np.array_equal(d.to_host(count), want)
            d.copy_from_host(np.zeros(2 * count, dtype=np.uint64))
            ba.ops.get_powers_by_ref_bf(b.sub(0, 1), off, bitrev, d.sub(0, count))
            ba.synchronize()
            assert np.array_equal(d.to_host(count), want)
            if count <= 4000:  # the oracle's per-element extension pow is slow in bulk
                want_e = gl_oracle.powers_by_val_ef(e0, e1, off, bitrev, count)
                ba.ops.get_powers_by_val_ef((e0, e1), off, bitrev, d)
                ba.synchronize()
                assert np.array_equal(d.to_host(), want_e)
                d.copy_from_host(np.zeros(2 * count, dtype=np.uint64))
                ba.ops.get_powers_by_ref_ef(be, off, bitrev, d)
                ba.synchronize()
                assert np.array_equal(d.to_host(), want_e)
        finally:
            for t in (d, b, be):
                t.free()


def test_omega_shift(ba, tables):
    rng = np.random.default_rng(9)
    for ld, off, inv, shift, count in [(10, 0, False, 1, 1024), (16, 99, True, 3, 50000),
                                        (30, 2**32 - 7, False, 2**31 + 1, 5000), (0, 5, False, 7, 9)]:
        log(f"omega_shift log_degree={ld} offset={off} inverse={inv} shift={shift} count={count}")
        x = raw_u64(rng, count)
        want = gl_oracle.omega_shift(x, ld, off, inv, shift)
        src = ba.DeviceAllocation.from_host(x)
        dst = ba.DeviceAllocation(count)
        try:
            ba.ops.omega_shift(src, ld, off, inv, shift, dst)
            ba.synchronize()
            assert np.array_equal(dst.to_host(), want)
            ba.ops.omega_shift_in_place(src, ld, off, inv, shift)
            ba.synchronize()
            assert np.array_equal(src.to_host(), want)
        finally:
            src.free()
            dst.free()
    a = ba.DeviceAllocation(100)
    try:
        with pytest.raises(ba.BoojumAmdError):  # partial overlap
            ba.check(ba.lib().ba_omega_shift(a.ptr, 5, 0, 0, 1, a.ptr + 8, 50, None))
        with pytest.raises(ba.BoojumAmdError):
            ba.ops.get_powers_of_w(31, 0, False, False, a)
    finally:
        a.free()


def test_diag_butterfly_rate_output(ba):
    """ba_diag_butterfly_rate's word per thread equals the same butterfly loop on Python integers
    (v[c], v[c+1] <- v[c] + w v[c+1], v[c] - w v[c+1], 8 values, xor of the final values)."""

    blocks, iters = 1, 3
    rng = np.random.default_rng(31)
    x = raw_u64(rng, 256 * blocks)
    d_in, d_out = ba.DeviceAllocation.from_host(x), ba.DeviceAllocation(256 * blocks)
    try:
        ba.check(ba.lib().ba_diag_butterfly_rate(d_in.ptr, d_out.ptr, blocks, iters, None), "butterfly_rate")
        ba.synchronize()
        got = d_out.to_host()
    finally:
        d_in.free()
        d_out.free()
    for t in range(0, 256, 17):
        xi = int(x[t])
        w = (xi % P) | 1
        v = [((xi + c) % 2**64) % P for c in range(8)]
        for _ in range(iters):
            for c in range(8):
                a, b = v[c], v[(c + 1) % 8] * w % P
                v[c], v[(c + 1) % 8] = (a + b) % P, (a - b) % P
        acc = 0
        for c in range(8):
            acc ^= v[c]
        assert int(got[t]) == acc, t


def test_stepped_powers_past_the_grid_stride(ba, tables):
    """More elements than the launch has lanes (2^24 + 4097 > 65536 x 256), so lanes step their
    power by omega^(stride ...) more than once; offsets and shifts make the u32 exponent wrap.
    Sampled indices (lane starts, stride boundaries, the end) against big-integer closed forms."""
    OMEGA = 0x7E9BD009B86A0845  # omega_{2^30}
    count = (1 << 24) + 4097
    stride = 65536 * 256
    idx = sorted({0, 1, 4095, 4096, 4097, stride - 1, stride, stride + 1, stride + 4096, count - 1}
                 | set(int(v) for v in np.random.default_rng(41).integers(0, count, 300)))
    d = ba.DeviceAllocation(count)
    src_host = raw_u64(np.random.default_rng(42), count)
    d_src = ba.DeviceAllocation.from_host(src_host)
    try:
        for log_degree, offset, inverse in ((20, 0xFFFFFC18, False), (24, 12345, True), (30, 7, False)):
            ba.ops.get_powers_of_w(log_degree, offset, inverse, False, d)
            ba.synchronize()
            got = d.to_host()
            for i in idx:
                e = ((((i + offset) % 2**32) << (30 - log_degree)) % 2**32) % 2**30
                want = pow(OMEGA, (-e) % 2**30 if inverse else e, P)
                assert int(got[i]) == want, (log_degree, offset, inverse, i)
        for log_degree, offset, inverse, shift in ((24, 0xFFFFF000, False, 3), (20, 5, True, 0x10001)):
            ba.ops.omega_shift(d_src, log_degree, offset, inverse, shift, d)
            ba.synchronize()
            got = d.to_host()
            for i in idx:
                e = (((shift * ((i + offset) % 2**32)) % 2**32) << (30 - log_degree)) % 2**32 % 2**30
                w = pow(OMEGA, (-e) % 2**30 if inverse else e, P)
                assert int(got[i]) == int(src_host[i]) % P * w % P, (log_degree, offset, inverse, shift, i)
    finally:
        d.free()
        d_src.free()
