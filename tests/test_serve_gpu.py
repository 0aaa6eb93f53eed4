"""GPU tests of serve mode (ba_server_*): every request's result is the same when it runs alone
and when it runs in a mix of concurrently submitted requests, and matches the oracle; admission
keeps the arena within its budget; invalid requests are refused."""
import threading

import numpy as np
import pytest

import gl_oracle
from gpu_util import column_seed, log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    ctx.destroy()


def make_requests(ba, rng, count):
    reqs = []
    for i in range(count):
        log_n = int(rng.choice([3, 5, 10, 12, 13, 16]))
        cols = int(rng.integers(1, 6))
        lb = int(rng.choice([1, 2, 3]))
        n = 1 << log_n
        kind = ("host", "device", "synthetic")[i % 3]
        r = dict(log_n=log_n, cols=cols, lb=lb, kind=kind, seed=900 + i)
        if kind == "synthetic":
            r["x"] = np.concatenate([gl_oracle.fill_uniform(n, column_seed(r["seed"], k)) for k in range(cols)])
        else:
            stride = n + (i % 2) * 3
            r["stride"] = stride
            r["x"] = raw_u64(rng, cols * stride)
            if kind == "device":
                r["dev"] = ba.DeviceAllocation.from_host(r["x"])
        reqs.append(r)
    return reqs


def submit(srv, r):
    if r["kind"] == "synthetic":
        return srv.submit(r["log_n"], r["cols"], r["lb"], seed=r["seed"])
    src = r["x"] if r["kind"] == "host" else r["dev"]
    return srv.submit(r["log_n"], r["cols"], r["lb"], evals=src, in_stride=r["stride"])


def fetch(srv, r, res):
    assert res["status"] == 0, res
    out = srv.output(res, r["log_n"], r["cols"], r["lb"]).to_host()
    srv.release(res["ticket"])
    return out


def test_alone_equals_mixed_and_oracle(ba):
    rng = np.random.default_rng(11)
    reqs = make_requests(ba, rng, 18)
    srv = ba.serve.Server(ba._test_ctx, 1 << 30, max_inflight_batches=2)
    try:
        alone = []
        for r in reqs:  # each request alone: submit, wait, fetch
            alone.append(fetch(srv, r, srv.wait(submit(srv, r))))
        # the same requests submitted together from three client threads
        tickets = [None] * len(reqs)

        def client(idx):
            for i in idx:
                tickets[i] = submit(srv, reqs[i])

        th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 3)),)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        results = [srv.wait(t) for t in tickets]
        batches = {res["batch"] for res in results}
        log(f"mixed: {len(reqs)} requests in {len(batches)} batches; stats {srv.stats()}")
        for r, res, a in zip(reqs, results, alone):
            got = fetch(srv, r, res)
            assert np.array_equal(got, a), (r["log_n"], r["cols"], r["lb"], r["kind"])
            stride = r.get("stride", 1 << r["log_n"])
            want = gl_oracle.lde(r["x"], r["log_n"], r["cols"], r["lb"], in_stride=stride, nthreads=8)
            assert np.array_equal(got, want), (r["log_n"], r["cols"], r["lb"], r["kind"])
            if r["kind"] == "device":  # device inputs are read, not modified
                assert np.array_equal(r["dev"].to_host(), r["x"])
        st = srv.stats()
        assert st["completed"] == 2 * len(reqs) and st["failed"] == 0
        assert st["in_use_bytes"] == 0 and st["peak_bytes"] <= st["budget_bytes"]
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        assert st["batches"] < 2 * len(reqs)  # the concurrent submissions were batched together
    finally:
        srv.destroy()
        for r in reqs:
            if "dev" in r:
                r["dev"].free()


def test_admission_within_budget(ba):
    """A budget that holds about two requests at a time: every request completes, the arena
    never exceeds the budget, and a request larger than the budget is refused at submit."""
    log_n, cols, lb = 16, 4, 2
    per_req = (cols << log_n) * 8 * (1 + (1 << lb))
    budget = 2 * per_req + 4096
    srv = ba.serve.Server(ba._test_ctx, budget, max_inflight_batches=2)
    try:
        tickets = [srv.submit(log_n, cols, lb, seed=50 + i) for i in range(10)]
        done = {}
        for t in tickets:  # wait in order and release, which lets the next requests in
            res = srv.wait(t)
            assert res["status"] == 0
            done[t] = srv.output(res, log_n, cols, lb).to_host()
            srv.release(t)
        st = srv.stats()
        log(f"budget {budget}: {st}")
        assert st["completed"] == 10 and st["peak_bytes"] <= budget
        # request 3 against the oracle
        x = np.concatenate([gl_oracle.fill_uniform(1 << log_n, column_seed(53, k)) for k in range(cols)])
        assert np.array_equal(done[tickets[3]], gl_oracle.lde(x, log_n, cols, lb, nthreads=8))
        with pytest.raises(ba.BoojumAmdError) as e:
            srv.submit(18, 4, 2, seed=1)
        assert e.value.code == 2  # BA_ERROR_OUT_OF_MEMORY
    finally:
        srv.destroy()


def test_bad_requests(ba):
    srv = ba.serve.Server(ba._test_ctx, 1 << 24)
    try:
        for log_n, cols, lb in [(0, 1, 1), (10, 0, 1), (10, 1, 0), (29, 1, 1)]:
            with pytest.raises(ba.BoojumAmdError):
                srv.submit(log_n, cols, lb, seed=1)
        with pytest.raises(ba.BoojumAmdError):
            srv.release(12345)
        with pytest.raises(ba.BoojumAmdError):
            srv.poll(12345)
    finally:
        srv.destroy()
    with pytest.raises(ba.BoojumAmdError):
        ba.serve.Server(ba._test_ctx, 1 << 24, max_inflight_batches=0)


def test_config_shapes_alone_equals_mixed(ba):
    """Requests of the serve configuration's shapes (8..16 columns, log_n up to 21, blowup 2, 4
    and 8, synthetic inputs): each result is bit-identical alone and when all are submitted at
    once, and a sampled column of each matches the oracle."""
    shapes = [(12, 16, 3), (21, 8, 1), (18, 12, 2), (14, 9, 3), (20, 10, 2), (16, 16, 1), (19, 8, 3)]
    srv = ba.serve.Server(ba._test_ctx, 24 << 30, max_inflight_batches=2)
    try:
        alone = []
        for i, (log_n, cols, lb) in enumerate(shapes):
            res = srv.wait(srv.submit(log_n, cols, lb, seed=7000 + i))
            assert res["status"] == 0
            alone.append(srv.output(res, log_n, cols, lb).to_host())
            srv.release(res["ticket"])
        tickets = [srv.submit(log_n, cols, lb, seed=7000 + i) for i, (log_n, cols, lb) in enumerate(shapes)]
        for i, ((log_n, cols, lb), t) in enumerate(zip(shapes, tickets)):
            res = srv.wait(t)
            assert res["status"] == 0
            got = srv.output(res, log_n, cols, lb).to_host()
            srv.release(t)
            log(f"serve shape log_n={log_n} cols={cols} blowup={1 << lb}: batch {res['batch']}")
            assert np.array_equal(got, alone[i]), (log_n, cols, lb)
            n, k = 1 << log_n, cols - 1
            x = gl_oracle.fill_uniform(n, column_seed(7000 + i, k))
            want = gl_oracle.lde(x, log_n, 1, lb, nthreads=8)
            assert np.array_equal(got[k * (n << lb):(k + 1) * (n << lb)], want), (log_n, cols, lb)
        st = srv.stats()
        log(f"stats {st}")
        assert st["failed"] == 0 and st["in_use_bytes"] == 0
    finally:
        srv.destroy()


def test_full_shapes_poisson_clients(ba):
    """Requests at the serve configuration's full shapes: log_n 22 and 23, 128 and 256 columns,
    blowup 8, submitted by three client threads at Poisson-timed arrivals while earlier ones run.
    Each result equals the same request run alone (first, middle and last column compared) and
    one column of each matches the oracle."""
    import time
    shapes = [(22, 16, 1), (23, 8, 2), (16, 128, 3), (12, 256, 3), (20, 12, 3), (14, 9, 2), (18, 40, 1),
              (13, 200, 2), (21, 10, 1)]
    srv = ba.serve.Server(ba._test_ctx, 24 << 30, max_inflight_batches=2)

    def sample_cols(res, log_n, cols, lb):
        out = srv.output(res, log_n, cols, lb)
        m = 1 << (log_n + lb)
        return {k: out.sub(k * m, m).to_host() for k in sorted({0, cols // 2, cols - 1})}

    try:
        alone = []
        for i, (log_n, cols, lb) in enumerate(shapes):
            log(f"alone log_n={log_n} cols={cols} blowup={1 << lb}")
            res = srv.wait(srv.submit(log_n, cols, lb, seed=8100 + i))
            assert res["status"] == 0
            alone.append(sample_cols(res, log_n, cols, lb))
            srv.release(res["ticket"])
        tickets = [None] * len(shapes)
        rng = np.random.default_rng(81)
        delays = rng.exponential(0.004, len(shapes))

        def client(idx):
            for i in idx:
                time.sleep(float(delays[i]))
                log_n, cols, lb = shapes[i]
                tickets[i] = srv.submit(log_n, cols, lb, seed=8100 + i)

        threads = [threading.Thread(target=client, args=(list(range(c, len(shapes), 3)),)) for c in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        batches = set()
        for i, (log_n, cols, lb) in enumerate(shapes):
            res = srv.wait(tickets[i])
            assert res["status"] == 0, res
            batches.add(res["batch"])
            got = sample_cols(res, log_n, cols, lb)
            srv.release(tickets[i])
            log(f"mixed log_n={log_n} cols={cols} blowup={1 << lb}: batch {res['batch']}, "
                f"queue {res['queue_ms']:.1f} ms, latency {res['latency_ms']:.1f} ms")
            for k in got:
                assert np.array_equal(got[k], alone[i][k]), (log_n, cols, lb, k)
            k = cols - 1
            x = gl_oracle.fill_uniform(1 << log_n, column_seed(8100 + i, k))
            assert np.array_equal(got[k], gl_oracle.lde(x, log_n, 1, lb, nthreads=8)), (log_n, cols, lb)
        st = srv.stats()
        log(f"stats {st}, batches {sorted(batches)}")
        assert st["failed"] == 0 and st["in_use_bytes"] == 0
    finally:
        srv.destroy()


def test_bypass_admits_small_requests_past_a_waiting_head(ba):
    """A budget that holds one large request: while a second large request waits for memory (the
    first one's outputs are still held), small requests submitted after it are admitted and
    complete (max_bypass 32); with max_bypass 0 they stay queued behind it until the first large
    result is released.  Every small result matches the oracle on one column."""
    import time
    big, small = (16, 16, 2), (10, 2, 1)
    per = lambda s: (s[1] << s[0]) * 8 * (1 + (1 << s[2]))  # noqa: E731
    budget = per(big) + 6 * per(small) + 8192

    def poll_until(srv, t, seconds):
        end = time.perf_counter() + seconds
        while time.perf_counter() < end:
            r = srv.poll(t)
            if r is not None:
                return r
            time.sleep(0.002)
        return None

    for bypass in (32, 0):
        srv = ba.serve.Server(ba._test_ctx, budget, max_inflight_batches=2)
        srv.set_max_bypass(bypass)
        try:
            t_big1 = srv.submit(*big, seed=1)
            assert srv.wait(t_big1)["status"] == 0  # its outputs stay held until released
            t_big2 = srv.submit(*big, seed=2)  # waits for memory
            t_small = [srv.submit(*small, seed=10 + i) for i in range(4)]
            early = [poll_until(srv, t, 10.0 if bypass else 0.2) for t in t_small]
            assert srv.poll(t_big2) is None  # still waiting for the first one's outputs
            if bypass:
                assert all(r is not None and r["status"] == 0 for r in early), early
            else:
                assert all(r is None for r in early), early
            srv.release(t_big1)
            assert srv.wait(t_big2)["status"] == 0
            smalls = early if bypass else [srv.wait(t) for t in t_small]
            for i, (r, t) in enumerate(zip(smalls, t_small)):
                assert r["status"] == 0
                got = srv.output(r, *small).sub(0, 1 << (small[0] + small[2])).to_host()
                x = gl_oracle.fill_uniform(1 << small[0], column_seed(10 + i, 0))
                assert np.array_equal(got, gl_oracle.lde(x, small[0], 1, small[2], nthreads=4))
                srv.release(t)
            srv.release(t_big2)
            log(f"bypass {bypass}: small queue ms {[round(r['queue_ms'], 2) for r in smalls]}")
        finally:
            srv.destroy()


def test_batches_hold_one_size_class(ba):
    """Requests submitted together but of different output size classes (below 2^26 elements,
    below 2^30, from 2^30 up) run in different batches, so the small ones do not wait for the
    large one merged with them; results match the oracle on sampled columns."""
    shapes = [(20, 64, 2)] + [(12, 8, 1)] * 5 + [(14, 4, 2)] * 2  # 2^28 output, then < 2^26 each
    srv = ba.serve.Server(ba._test_ctx, 8 << 30, max_inflight_batches=2)
    try:
        tickets = [srv.submit(log_n, cols, lb, seed=9300 + i) for i, (log_n, cols, lb) in enumerate(shapes)]
        results = [srv.wait(t) for t in tickets]
        classes = {}
        for (log_n, cols, lb), res in zip(shapes, results):
            assert res["status"] == 0, res
            e = cols << (log_n + lb)
            classes.setdefault(0 if e < 1 << 26 else (1 if e < 1 << 30 else 2), set()).add(res["batch"])
        log(f"batches by size class: {classes}")
        assert len(classes) == 2 and not (classes[0] & classes[1])
        for i in (0, 1, len(shapes) - 1):
            log_n, cols, lb = shapes[i]
            m = 1 << (log_n + lb)
            k = cols - 1
            got = srv.output(results[i], log_n, cols, lb).sub(k * m, m).to_host()
            x = gl_oracle.fill_uniform(1 << log_n, column_seed(9300 + i, k))
            assert np.array_equal(got, gl_oracle.lde(x, log_n, 1, lb, nthreads=8)), shapes[i]
        for t in tickets:
            srv.release(t)
        st = srv.stats()
        assert st["failed"] == 0 and st["in_use_bytes"] == 0
    finally:
        srv.destroy()
