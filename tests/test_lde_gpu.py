"""GPU parity of the batched coset LDE (ba_lde_batch) against the oracle's restatement of
shivini's composition (crates/shivini/src/data_structures/storage.rs:154-260)."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import P, bitrev_perm, log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    ctx.destroy()


def run_lde(ba, evals_host, log_n, cols, log_blowup, layout="column", keep_evals=False, in_stride=None):
    n = 1 << log_n
    B = 1 << log_blowup
    in_stride = n if in_stride is None else in_stride
    d_ev = ba.DeviceAllocation.from_host(evals_host)
    if layout == "column":
        col_stride, coset_stride = B * n, n
    else:  # coset-major, as shivini stores each coset as its own matrix
        col_stride, coset_stride = n, cols * n
    d_out = ba.DeviceAllocation(B * n * cols)
    d_mono = ba.DeviceAllocation(n * cols) if keep_evals else None
    try:
        ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, log_blowup, d_out, in_stride=in_stride,
                         monomials=d_mono, out_col_stride=col_stride, out_coset_stride=coset_stride)
        ba.synchronize()
        out = d_out.to_host()
        ev_after = d_ev.to_host()
        mono = d_mono.to_host() if d_mono is not None else None
    finally:
        d_ev.free()
        d_out.free()
        if d_mono is not None:
            d_mono.free()
    # normalise to column-contiguous layout
    if layout != "column":
        out = np.concatenate([np.concatenate([out[c * cols * n + k * n: c * cols * n + (k + 1) * n]
                                              for c in range(B)]) for k in range(cols)])
    return out, ev_after, mono


@pytest.mark.parametrize("log_blowup", [1, 2, 3])
def test_lde_small_sizes(ba, log_blowup):
    for log_n in list(range(1, 15)) + [16]:
        cols = 3
        n = 1 << log_n
        log(f"lde parity log_n={log_n} blowup=2^{log_blowup}")
        ev = raw_u64(np.random.default_rng(log_n * 7 + log_blowup), cols * n)
        want = gl_oracle.lde(ev, log_n, cols, log_blowup, nthreads=4)
        got, ev_after, _ = run_lde(ba, ev, log_n, cols, log_blowup)
        assert np.array_equal(got, want), (log_n, log_blowup)
        # the evaluations were replaced by the monomials in bit-reversed order
        mono_br = gl_oracle.batch_ntt(ev, log_n, cols, n, n, False, True)
        assert np.array_equal(ev_after, mono_br)


@pytest.mark.parametrize("layout", ["column", "coset"])
def test_lde_layouts_and_scratch(ba, layout):
    for log_n in (5, 12, 13, 17):
        cols, lb = 5, 2
        n = 1 << log_n
        log(f"lde layout={layout} log_n={log_n}")
        ev = raw_u64(np.random.default_rng(log_n), cols * (n + 3))
        want = gl_oracle.lde(ev, log_n, cols, lb, in_stride=n + 3, nthreads=4)
        got, ev_after, mono = run_lde(ba, ev, log_n, cols, lb, layout=layout, keep_evals=True, in_stride=n + 3)
        assert np.array_equal(got, want)
        assert np.array_equal(ev_after, ev)  # evaluations kept when a monomial buffer is given
        assert np.array_equal(mono, gl_oracle.batch_ntt(ev, log_n, cols, n + 3, n, False, True))


@pytest.mark.parametrize("keep_evals", [False, True])
def test_lde_natural_monomials(ba, keep_evals):
    """ba_lde_batch_ex with BA_LDE_NATURAL_MONOMIALS: the same coset evaluations, and the
    monomials in natural order (shivini's into_monomials after its bit-reverse epilogue,
    crates/shivini/src/data_structures/storage.rs:155), in a separate buffer or in place."""
    for log_n in (1, 4, 9, 10, 13, 16):
        cols, lb = 3, 2
        n = 1 << log_n
        log(f"lde natural monomials log_n={log_n} keep_evals={keep_evals}")
        ev = raw_u64(np.random.default_rng(100 + log_n), cols * n)
        d_ev = ba.DeviceAllocation.from_host(ev)
        d_out = ba.DeviceAllocation((n << lb) * cols)
        d_mono = ba.DeviceAllocation(n * cols) if keep_evals else None
        try:
            ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono,
                             natural_monomials=True)
            ba.synchronize()
            got = d_out.to_host()
            mono = (d_mono if keep_evals else d_ev).to_host()
        finally:
            d_ev.free()
            d_out.free()
            if d_mono is not None:
                d_mono.free()
        assert np.array_equal(got, gl_oracle.lde(ev, log_n, cols, lb, nthreads=4)), log_n
        mono_br = gl_oracle.batch_ntt(ev, log_n, cols, n, n, False, True)
        perm = bitrev_perm(log_n)
        want = np.concatenate([mono_br[k * n:(k + 1) * n][perm] for k in range(cols)])
        assert np.array_equal(mono, want), log_n
        # natural monomials are the inverse NTT in natural order: evaluating them at omega^i
        # (a forward natural -> bit-reversed NTT) gives back the evaluations, bit-reversed
        back = gl_oracle.batch_ntt(mono, log_n, cols, n, n, False, False)
        ev_red = ev % np.uint64(P)
        assert np.array_equal(np.concatenate([back[k * n:(k + 1) * n][perm] for k in range(cols)]), ev_red)


def test_lde_golden(ba, golden):
    for log_n in range(1, 8):
        ev = golden[f"lde_in_{log_n}"]
        for lb in (1, 2, 3):
            got, _, _ = run_lde(ba, ev, log_n, 1, lb)
            assert np.array_equal(got, golden[f"lde_{log_n}_{lb}"])


def test_lde_headline_size_samples(ba):
    """64 columns of 2^20 at blowup 2 (the benchmark's shape): two columns checked against the
    oracle, every column through the interpolation round trip."""
    log_n, cols, lb = 20, 64, 1
    n, B = 1 << log_n, 2
    d_ev = ba.DeviceAllocation(n * cols)
    d_keep = ba.DeviceAllocation(n * cols)
    d_out = ba.DeviceAllocation(B * n * cols)
    d_tmp = ba.DeviceAllocation(n * cols)
    try:
        ba.runtime.fill_uniform(d_ev, n, cols, 2024)
        d_keep.copy_from_device(d_ev)
        ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out)
        ba.synchronize()
        ev = d_keep.to_host()
        out = d_out.to_host()
        for k in (0, 37):
            want = gl_oracle.lde(ev[k * n:(k + 1) * n], log_n, 1, lb, nthreads=8)
            assert np.array_equal(out[k * B * n:(k + 1) * B * n], want), k
        # every coset interpolates back to the evaluations:
        # b2n inverse with the coset factors -> natural monomials -> n2b forward -> bitrev evals
        perm = bitrev_perm(log_n)
        for c in range(B):
            coset_idx = c  # bitrev over one bit is the identity
            for k in range(cols):
                d_tmp.sub(k * n, n).copy_from_device(d_out.sub(k * B * n + c * n, n))
            ba.ntt.batch_ntt_internal(d_tmp.ptr, d_tmp.ptr, log_n, cols, n, n, True, True, lb, coset_idx)
            ba.ntt.batch_ntt_internal(d_tmp.ptr, d_tmp.ptr, log_n, cols, n, n, False, False, 0, 0)
            ba.synchronize()
            back = d_tmp.to_host()
            for k in range(cols):
                assert np.array_equal(back[k * n:(k + 1) * n][perm], ev[k * n:(k + 1) * n]), (c, k)
    finally:
        for t in (d_ev, d_keep, d_out, d_tmp):
            t.free()


def test_lde_large_trace_samples(ba):
    """256 columns of 2^22 at blowup 4 (2^32 output elements, 34.4 GB): two columns checked
    against the oracle, eight columns of every coset through the interpolation round trip."""
    log_n, cols, lb = 22, 256, 2
    n, B = 1 << log_n, 1 << lb
    sample = [0, 1, 77, 128, 200, 254, 255, 31]
    d_ev = ba.DeviceAllocation(n * cols)
    d_out = ba.DeviceAllocation(B * n * cols)
    d_mono = ba.DeviceAllocation(n * cols)
    d_tmp = ba.DeviceAllocation(n * len(sample))
    try:
        ba.runtime.fill_uniform(d_ev, n, cols, 4242)
        log("large trace LDE 256 x 2^22 blowup 4")
        ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono)
        ba.synchronize()
        ev = {k: d_ev.sub(k * n, n).to_host() for k in sample}
        for k in (1, 254):
            want = gl_oracle.lde(ev[k], log_n, 1, lb, nthreads=8)
            assert np.array_equal(d_out.sub(k * B * n, B * n).to_host(), want), k
        perm = bitrev_perm(log_n)
        for c in range(B):
            coset_idx = int(format(c, f"0{lb}b")[::-1], 2)
            for i, k in enumerate(sample):
                d_tmp.sub(i * n, n).copy_from_device(d_out.sub(k * B * n + c * n, n))
            m = len(sample)
            ba.ntt.batch_ntt_internal(d_tmp.ptr, d_tmp.ptr, log_n, m, n, n, True, True, lb, coset_idx)
            ba.ntt.batch_ntt_internal(d_tmp.ptr, d_tmp.ptr, log_n, m, n, n, False, False, 0, 0)
            ba.synchronize()
            back = d_tmp.to_host()
            for i, k in enumerate(sample):
                assert np.array_equal(back[i * n:(i + 1) * n][perm], ev[k]), (c, k)
    finally:
        for t in (d_ev, d_out, d_mono, d_tmp):
            t.free()


def test_lde_rejects_bad_arguments(ba):
    d = ba.DeviceAllocation(1 << 10)
    o = ba.DeviceAllocation(1 << 12)
    try:
        for log_n, lb in [(0, 1), (5, 0), (28, 3)]:
            with pytest.raises(ba.BoojumAmdError):
                ba.lde.lde_batch(ba._test_ctx, d, log_n, 1, lb, o)
        with pytest.raises(ba.BoojumAmdError):
            ba.lde.lde_batch(ba._test_ctx, d, 5, 1, 1, d)  # output aliases evaluations
        with pytest.raises(ba.BoojumAmdError):  # coset 1 of column 0 is column 1's coset 0
            ba.lde.lde_batch(ba._test_ctx, d, 5, 2, 1, o, out_col_stride=32, out_coset_stride=32)
        with pytest.raises(ba.BoojumAmdError):  # output overlaps the evaluations' tail
            ba.lde.lde_batch(ba._test_ctx, d, 5, 2, 1, d.sub(40, 900))
        with pytest.raises(ba.BoojumAmdError):  # monomials overlap the evaluations
            ba.lde.lde_batch(ba._test_ctx, d, 5, 2, 1, o, monomials=d.sub(16, 64))
    finally:
        d.free()
        o.free()


def test_context_outlives_its_users(ba):
    """A context is not destroyed while a batch plan or a server still reads its tables."""
    ctx = ba.Context.create(12, 12)
    d = ba.DeviceAllocation(1 << 12)
    try:
        plan = ba.batch.BatchPlan(ctx, [ba.batch.ntt_job(d, d, 12, 1, 1 << 12, 1 << 12)])
        with pytest.raises(ba.BoojumAmdError):
            ctx.destroy()
        assert ctx.handle.value  # a refused destroy leaves the context usable
        plan.launch()
        ba.synchronize()
        plan.destroy()
        srv = ba.serve.Server(ctx, 1 << 24)
        with pytest.raises(ba.BoojumAmdError):
            ctx.destroy()
        srv.destroy()
    finally:
        d.free()
        ctx.destroy()


@pytest.mark.parametrize("lb", [2, 4])
def test_lde_graph_replay(ba, lb):
    """The LDE captured as a hipGraph (as the benchmark runs it) gives the eager result on every
    replay, with eager work and synchronisation between replays.  Blowup 4 forks cosets 1..3 onto
    three auxiliary streams; blowup 16 puts cosets 1..15 on all seven, two or three per stream."""
    import ctypes

    log_n, cols = 16, 6
    n, B = 1 << log_n, 1 << lb
    rng = np.random.default_rng(31)
    ev = raw_u64(rng, n * cols)
    d_ev, d_mono = ba.DeviceAllocation.from_host(ev), ba.DeviceAllocation(n * cols)
    d_out, d_eager = ba.DeviceAllocation(B * n * cols), ba.DeviceAllocation(B * n * cols)
    stream = ba.CudaStream()
    lib = ba.lib()
    try:
        ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_eager, monomials=d_mono)
        ba.synchronize()
        eager = d_eager.to_host()
        assert np.array_equal(eager, gl_oracle.lde(ev, log_n, cols, lb, nthreads=8))
        ba.check(lib.ba_stream_begin_capture(stream.handle), "capture")
        ba.lde.lde_batch(ba._test_ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono, stream=stream)
        graph = ctypes.c_void_p(None)
        ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(graph)), "end capture")
        try:
            for rep in range(4):
                log(f"lde graph replay {rep}")
                d_out.copy_from_host(np.zeros(B * n * cols, dtype=np.uint64))  # eager work between replays
                ba.check(lib.ba_graph_launch(graph, stream.handle), "graph launch")
                stream.synchronize()
                assert np.array_equal(d_out.to_host(), eager), rep
                assert np.array_equal(d_ev.to_host(), ev)  # evaluations kept (monomial buffer given)
        finally:
            ba.check(lib.ba_graph_destroy(graph), "graph destroy")
    finally:
        for t in (d_ev, d_mono, d_out, d_eager):
            t.free()
        stream.destroy()
