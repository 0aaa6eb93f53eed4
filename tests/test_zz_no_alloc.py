"""ba_barycentric_batch_eval_noalloc (the entry point that allocates nothing: partial sums in the
library's per-GPU array) against ba_barycentric_batch_eval with a caller workspace, bit-exact,
on two calls in a row (the second smaller, so stale partial sums of the first would show), for
all three kinds; and its size limit."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import P, log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    return boojum_amd


def _eval(ba, kind, ys, count, num_polys, coeffs, noalloc):
    out_n = num_polys * (1 if kind == 0 else 2)
    d_ys, d_l, d_e = ba.DeviceAllocation.from_host(ys), ba.DeviceAllocation.from_host(coeffs), ba.DeviceAllocation(out_n)
    d_t = None
    try:
        if noalloc:
            ba.barycentric.batch_eval_noalloc(kind, d_ys, count, num_polys, d_l, d_e)
        else:
            d_t = ba.DeviceAllocation(ba.barycentric.batch_eval_temp_elems(kind, count, num_polys))
            ba.barycentric.batch_eval(kind, d_ys, count, num_polys, d_l, d_t, d_e)
        ba.synchronize()
        return d_e.to_host()
    finally:
        for t in (d_ys, d_l, d_e, d_t):
            if t is not None:
                t.free()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_noalloc_matches_workspace_entry_point(ba, kind):
    rng = np.random.default_rng(40 + kind)
    ext = kind != 0
    # the first call fills most of the array (2^18 partial sums for kind 0), the second is smaller
    for log_count, num_polys in [(20, 1024 if kind == 0 else 512), (13, 7)]:
        count = 1 << log_count
        x = (int(rng.integers(0, 2**64, dtype=np.uint64)), int(rng.integers(0, 2**64, dtype=np.uint64))) if ext \
            else int(rng.integers(0, 2**64, dtype=np.uint64))
        xo = (x[0] % P, x[1] % P) if ext else x % P
        coeffs = gl_oracle.barycentric_coeffs(xo, 7, count, True, ext=ext)
        ys = raw_u64(rng, count * num_polys * (2 if kind == 2 else 1))
        log(f"noalloc kind={kind} log_count={log_count} polys={num_polys}")
        want = _eval(ba, kind, ys, count, num_polys, coeffs, noalloc=False)
        got = _eval(ba, kind, ys, count, num_polys, coeffs, noalloc=True)
        assert np.array_equal(got, want), (kind, log_count, num_polys)
        if log_count <= 13:
            assert np.array_equal(got, gl_oracle.barycentric_eval(kind, ys, count, num_polys, coeffs))


def test_noalloc_rejects_requests_beyond_its_array(ba):
    # 2^24 rows x 256 columns needs 2^20 partial sums: beyond the 2^19-element array
    assert ba.barycentric.batch_eval_temp_elems(0, 1 << 24, 256) > ba.barycentric.NOALLOC_TEMP_ELEMS
    d = ba.DeviceAllocation(4)
    try:
        with pytest.raises(ba.BoojumAmdError):
            ba.barycentric.batch_eval_noalloc(0, d, 1 << 24, 256, d, d)
    finally:
        d.free()
