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
