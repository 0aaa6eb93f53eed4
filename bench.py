#!/usr/bin/env python3
"""Benchmark: batched Goldilocks coset LDE on one MI355X.

Workload (the headline configuration): 64 columns of 2^20 uniform field elements, interpolated
and evaluated on 2 cosets of the blowup-2 domain with shift 7 (inverse NTT, coset scale, forward
NTT, bit-reversed output: 64 x 2^21 output elements per step), the whole LDE replayed as one
captured hipGraph.  Inputs are generated on the device and resident in HBM before timing.

Printed (one JSON line): Melem/s of output, the batched-vs-per-column-call speedup, the
roofline of the dominant kernel (timed per launch with HIP events through the library's launch
tracing), and the CPU baseline (the oracle's restatement of boojum's CPU NTT on a bounded sample
of columns, on this host's cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zksync-crypto-gpu_amd"))

HBM_PEAK_GBS = 8000.0
VALU_PEAK_WINST = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second on MI355X
# HBM bytes per launch of each kernel on this workload, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs scaled by a copy kernel (profiles/lde_64x2p20_blowup2_passes.md)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lde_64x2p20_blowup2_traffic.json")
METRIC = "LDE Melem/s + % of HBM roofline, batched vs per-call; serve p50/p99 ms"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=30, help="untimed steps first: the clocks settle after ~30")
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--cols", type=int, default=64)
    ap.add_argument("--log-blowup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-per-call", action="store_true", help="skip the per-column-call leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--no-tune", action="store_true", help="uneven and serve legs: built-in pass plans, no autotuning")
    ap.add_argument("--tune-headline", action="store_true",
                    help="autotune the headline size class too (default: built-in plan, which the tuner "
                         "keeps for 2^20)")
    ap.add_argument("--no-large", action="store_true", help="skip the 256 x 2^22 blowup-4 leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serve-mode leg")
    ap.add_argument("--no-single", action="store_true", help="skip the single 2^16 NTT leg")
    ap.add_argument("--no-merkle", action="store_true", help="skip the Poseidon2 Merkle tree over the LDE output")
    ap.add_argument("--no-ops", action="store_true", help="skip omega_shift, powers and barycentric evaluation")
    ap.add_argument("--no-harness", action="store_true",
                    help="skip the leg with the reference benchmark's grid (2^28-element matrices)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a 2^27-element copy kernel (canonicalise, 8 B read + 8 B written "
                         "per element) that scales the FETCH_SIZE/WRITE_SIZE counters")
    ap.add_argument("--only-classes", default=None, metavar="LO:HI",
                    help="profiling aid: only forward natural-order batch NTTs of 2^26 elements per size "
                         "class log_n in LO..HI (2^(26-log_n) columns), --steps launches each")
    ap.add_argument("--only-single", action="store_true",
                    help="run only the single 2^16 NTT leg (one call per transform) and print its JSON line")
    ap.add_argument("--serial-cosets", action="store_true",
                    help="run every LDE's coset NTTs in order on the caller's stream (sets "
                         "BOOJUM_AMD_LDE_SERIAL=1), so that a profiler's per-kernel durations are each "
                         "launch's own rather than overlapping the other cosets'")
    return ap.parse_args()


def cpu_baseline(log_n, log_blowup):
    """The oracle (plain-C restatement of boojum's CPU NTT) on a bounded sample of columns."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gl_oracle

    # one column per thread at a time, on the host cores this process may use, at most 16: a
    # one-GPU box's share of its host is 16 cores, whatever the machine's total
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    threads = max(1, min(16, usable))
    cols = 2 * threads  # about 10-30 s of CPU work in all
    n = 1 << log_n
    ev = gl_oracle.fill_uniform(cols * n, 7)
    t0 = time.perf_counter()
    gl_oracle.lde(ev, log_n, cols, log_blowup, nthreads=threads)
    dt = time.perf_counter() - t0
    out_elems = cols * (n << log_blowup)
    # the single-column 2^16 forward NTT on one core
    x16 = gl_oracle.fill_uniform(1 << 16, 0)
    t1 = time.perf_counter()
    for _ in range(20):
        gl_oracle.batch_ntt(x16, 16, 1, 1 << 16, 1 << 16, False, False, nthreads=1)
    single_us = (time.perf_counter() - t1) / 20 * 1e6
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": round(out_elems / dt / 1e6, 3),
        "unit": "Melem/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "usable_cpus": usable,
        "cpu_model": model,
        "kind": "port",
        "sample": f"{cols} columns of the same 2^{log_n} blowup-{1 << log_blowup} LDE on {threads} threads, "
                  f"{dt:.2f} s wall ({dt * threads:.0f} thread-s)",
        "single_ntt_2p16_us_1core": round(single_us, 1),
    }


def time_calls(ba, ctx, stream, log_n, calls):
    """(us per call of back-to-back one-column forward NTT calls, HIP events around them; kernel
    launches per call, counted by the library's launch tracing on one untimed call)."""
    from boojum_amd.profile import LaunchTrace
    n = 1 << log_n
    d_in, d_out = ba.DeviceAllocation(n), ba.DeviceAllocation(n)
    try:
        ba.runtime.fill_uniform(d_in, n, 1, 0)
        for _ in range(10):
            ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, 1, n, n, False, False, stream=stream)
        stream.synchronize()
        e0, e1 = ba.CudaEvent(), ba.CudaEvent()
        e0.record(stream)
        for _ in range(calls):
            ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, 1, n, n, False, False, stream=stream)
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_ms(e1) / calls * 1e3
        e0.destroy()
        e1.destroy()
        with LaunchTrace(16) as tr:
            ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, 1, n, n, False, False, stream=stream)
        stream.synchronize()
        launches = tr.count
        tr.destroy()
    finally:
        d_in.free()
        d_out.free()
    return us, launches


def run_single(ba, ctx, stream, log_n=16, calls=200):
    """One column of 2^16, natural -> bit-reversed forward NTT, one library call per transform
    (the reference's CPU-runnable case), inputs resident in HBM.  Beside it, the same for one
    column of 2^8, which is one launch: the per-call floor a one-launch call pays."""
    us, launches = time_calls(ba, ctx, stream, log_n, calls)
    floor, floor_launches = time_calls(ba, ctx, stream, 8, calls)
    return {"workload": f"1 column x 2^{log_n} forward NTT, natural in, bit-reversed out, one call each",
            "us_per_call": round(us, 2), "melem_s": round((1 << log_n) / us, 1), "launches_per_call": launches,
            "floor": {"workload": "1 column x 2^8, the same call", "launches_per_call": floor_launches,
                      "us_per_call": round(floor, 2)}}


def uneven_jobs(ba, seed=3, jobs_n=512):
    """The uneven batch's jobs: single columns, log_n uniform in 12..22, random forward/inverse
    flag, 30% coset LDEs at blowup 2 or 4; job i's input is ba_fill_uniform(seed 1000 + i).
    Returns (plan jobs, per-job call tuples, device buffers, output elements)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    jobs, bufs, calls, out_elems = [], [], [], 0
    for i in range(jobs_n):
        log_n = int(rng.integers(12, 23))
        n = 1 << log_n
        if rng.random() < 0.3:
            lb = int(rng.integers(1, 3))
            d_in, d_mono, d_out = ba.DeviceAllocation(n), ba.DeviceAllocation(n), ba.DeviceAllocation(n << lb)
            ba.runtime.fill_uniform(d_in, n, 1, 1000 + i)
            jobs.append(ba.batch.lde_job(d_in, d_out, log_n, 1, lb, monomials=d_mono))
            calls.append(("lde", d_in, d_out, d_mono, log_n, lb))
            bufs += [d_in, d_mono, d_out]
            out_elems += n << lb
        else:
            inverse = bool(rng.integers(0, 2))
            d_in, d_out = ba.DeviceAllocation(n), ba.DeviceAllocation(n)
            ba.runtime.fill_uniform(d_in, n, 1, 1000 + i)
            jobs.append(ba.batch.ntt_job(d_in, d_out, log_n, 1, n, n, False, inverse))
            calls.append(("ntt", d_in, d_out, None, log_n, inverse))
            bufs += [d_in, d_out]
            out_elems += n
    return jobs, calls, bufs, out_elems


def run_uneven(ba, ctx, stream, seed=3, jobs_n=512, reps=5):
    """Uneven batch (uneven_jobs): one batch plan (grouped launches) replayed as a hipGraph,
    against one library call per job on the same stream."""
    import ctypes

    lib = ba.lib()
    jobs, calls, bufs, out_elems = uneven_jobs(ba, seed, jobs_n)
    plan = ba.batch.BatchPlan(ctx, jobs)
    ba.check(lib.ba_stream_begin_capture(stream.handle), "begin_capture")
    plan.launch(stream)
    graph = ctypes.c_void_p(None)
    ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(graph)), "end_capture")

    def batched():
        ba.check(lib.ba_graph_launch(graph, stream.handle), "graph_launch")

    def per_call():
        for kind, d_in, d_out, d_mono, log_n, x in calls:
            n = 1 << log_n
            if kind == "lde":
                ba.lde.lde_batch(ctx, d_in, log_n, 1, x, d_out, monomials=d_mono, stream=stream)
            else:
                ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, 1, n, n, False, x, stream=stream)

    def timed(fn):
        fn()
        stream.synchronize()
        e0, e1 = ba.CudaEvent(), ba.CudaEvent()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        e0.destroy()
        e1.destroy()
        return ms

    b_ms, p_ms = timed(batched), timed(per_call)
    st = plan.stats()
    ba.check(lib.ba_graph_destroy(graph), "graph_destroy")
    plan.destroy()
    for b in bufs:
        b.free()
    return {"jobs": jobs_n, "log_n": "12..22 uniform", "lde_fraction": 0.3, "output_elems": out_elems,
            "batched_ms": round(b_ms, 4), "batched_melem_s": round(out_elems / (b_ms / 1e3) / 1e6, 1),
            "per_call_ms": round(p_ms, 4), "per_call_melem_s": round(out_elems / (p_ms / 1e3) / 1e6, 1),
            "batched_speedup": round(p_ms / b_ms, 3), "plan_launches": st["launches"], "plan_levels": st["levels"]}


def butterfly_rate(ba, stream, blocks=4096, iters=64, reps=5):
    """Butterflies per second of ba_diag_butterfly_rate: the NTT passes' butterfly arithmetic
    (one general product, one canonical add and subtract) in registers, 2^20 threads x 64 x 8
    butterflies per launch, HIP events on the stream."""
    n = blocks * 256
    d_in, d_out = ba.DeviceAllocation(n), ba.DeviceAllocation(n)
    e0, e1 = ba.CudaEvent(), ba.CudaEvent()
    lib = ba.lib()
    try:
        ba.runtime.fill_uniform(d_in, n, 1, 77)

        def run():
            ba.check(lib.ba_diag_butterfly_rate(d_in.ptr, d_out.ptr, blocks, iters, stream.handle), "butterfly_rate")

        for _ in range(3):
            run()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
    finally:
        e0.destroy()
        e1.destroy()
        d_in.free()
        d_out.free()
    return n * iters * 8 / (ms / 1e3), ms


def run_harness(ba, stream, reps=10, samples=3):
    """The reference's own benchmark grid (crates/boojum-cuda/benches/ntt.rs:21-126): one
    2^28-element matrix, transformed out of place as 2^28 / 2^log_n columns by one
    batch_ntt_out_of_place call, for every (inverse, bitrev_inputs) mode; "batch_ntt" over log_n
    16..24, "batch_lde" over 16..22 with log_extension_degree 2 and coset_index 1.  Like its
    criterion harness, each sample times `reps` calls with events on the stream and reports time
    per individual NTT; here the median of `samples` samples, and the pass traffic (16 B per
    element) as a fraction of the HBM peak.  Inputs are uniform field elements (the reference
    feeds non-reduced u64, which the kernels read mod p at the same cost)."""
    problem = 1 << 28
    d_in, d_out = ba.DeviceAllocation(problem), ba.DeviceAllocation(problem)
    e0, e1 = ba.CudaEvent(), ba.CudaEvent()
    out = {"workload": "2^28-element matrix, out of place, 2^28 / 2^log_n columns per call",
           "unit": "us per NTT, fraction of HBM peak at 16 B per element"}
    try:
        ba.runtime.fill_uniform(d_in, problem, 1, 28)
        for group, log_ns, lext, coset in (("batch_ntt", range(16, 25), 0, 0), ("batch_lde", range(16, 23), 2, 1)):
            res = {}
            for inverse in (False, True):
                for bitrev in (False, True):
                    row = {}
                    for log_n in log_ns:
                        n = 1 << log_n
                        cols = problem >> log_n

                        def call():
                            ba.ntt.batch_ntt_out_of_place(d_in, d_out, log_n, cols, 0, 0, n, n, bitrev, inverse,
                                                          lext, coset, stream=stream)

                        call()
                        times = []
                        for _ in range(samples):
                            e0.record(stream)
                            for _ in range(reps):
                                call()
                            e1.record(stream)
                            e1.synchronize()
                            times.append(e0.elapsed_ms(e1) / reps)
                        ms = sorted(times)[len(times) // 2]
                        row[log_n] = [round(ms * 1e3 / cols, 3),
                                      round(16 * problem / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)]
                    res[("inverse" if inverse else "forward") + (", b2n" if bitrev else ", n2b")] = row
            out[group] = res
    finally:
        e0.destroy()
        e1.destroy()
        d_in.free()
        d_out.free()
    return out


def run_merkle(ba, ctx, stream, d_out, log_n, cols, lb, log_rows=1, reps=5):
    """Poseidon2 Merkle tree over the headline LDE's output (the coset evaluations the LDE just
    wrote, a cols x 2^(log_n + lb) column-major matrix): leaves of 2^log_rows rows across all
    columns, then every node layer up to the root, as GpuTreeHasher::build_merkle_tree
    (crates/boojum-cuda/src/poseidon2.rs:297-316) runs it.  Round constants are seeded field
    elements (boojum's table is not in the reference tree; the cost does not depend on them)."""
    import numpy as np

    rows_total = 1 << (log_n + lb)
    leaves = rows_total >> log_rows
    layers = (leaves.bit_length() - 1) + 1  # leaves up to the single root
    rng = np.random.default_rng(2023)
    ba.poseidon2.set_round_constants(ctx, rng.integers(0, 0xFFFFFFFF00000001, 360, dtype=np.uint64))
    d_tree = ba.DeviceAllocation(8 * leaves)
    values = d_out.sub(0, cols * rows_total)
    e0, e1 = ba.CudaEvent(), ba.CudaEvent()
    try:
        for _ in range(2):
            ba.poseidon2.build_merkle_tree(values, d_tree, log_rows, stream, layers)
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            ba.poseidon2.build_merkle_tree(values, d_tree, log_rows, stream, layers)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / reps
    finally:
        e0.destroy()
        e1.destroy()
        d_tree.free()
    hashed = cols * rows_total
    perms = leaves * -(-((cols << log_rows)) // 8) + (leaves - 1)
    return {"workload": f"Poseidon2 Merkle tree over the LDE output: {cols} columns x 2^{log_n + lb} rows, "
                        f"{1 << log_rows} rows per leaf, {layers} layers to the root",
            "ms": round(ms, 3), "hashed_melem_s": round(hashed / (ms / 1e3) / 1e6, 1),
            "permutations_per_s": round(perms / (ms / 1e3), 1), "permutations": perms,
            "input_GBs": round(8 * hashed / (ms / 1e3) / 1e9, 1)}


def run_ops(ba, stream, d_ev, d_tmp, log_n, cols, reps=5):
    """The table-sharing element-wise kernels and barycentric evaluation at the headline's size
    (cols x 2^log_n elements, resident in HBM), each against the HBM roofline: omega_shift (read +
    write, 16 B per element, ops_complex.rs:228-260), get_powers_of_w (write, 8 B per element,
    ops_complex.rs:31-52) and barycentric batch_eval of every column at one point (read, 8 B per
    element; the 2^log_n coefficients are re-read per column group, barycentric.rs:197-300)."""
    import numpy as np

    n, total = 1 << log_n, cols << log_n
    e0, e1 = ba.CudaEvent(), ba.CudaEvent()

    def timed(fn):
        for _ in range(2):
            fn()
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_ms(e1) / reps

    out = {}
    dst = d_tmp.sub(0, total)
    src = d_ev.sub(0, total)
    ms = timed(lambda: ba.ops.omega_shift(src, log_n, 0, False, 1, dst, stream))
    out["omega_shift"] = {"elems": total, "ms": round(ms, 4), "GBs": round(16 * total / ms / 1e6, 1),
                          "hbm_frac": round(16 * total / ms / 1e6 / HBM_PEAK_GBS, 4)}
    ms = timed(lambda: ba.ops.get_powers_of_w(26, 0, False, False, dst, stream))
    out["get_powers_of_w"] = {"elems": total, "ms": round(ms, 4), "GBs": round(8 * total / ms / 1e6, 1),
                              "hbm_frac": round(8 * total / ms / 1e6 / HBM_PEAK_GBS, 4)}
    x = ba.DeviceAllocation.from_host(np.array([0x1234567890ABCDEF % 0xFFFFFFFF00000001], dtype=np.uint64))
    common = ba.DeviceAllocation(1)
    coeffs = ba.DeviceAllocation(n)
    evals = ba.DeviceAllocation(cols)
    temp = ba.DeviceAllocation(max(1, ba.barycentric.batch_eval_temp_elems(ba.barycentric.BASE_AT_BASE, n, cols)))
    try:
        ba.barycentric.precompute_lagrange_coeffs(x, common, 7, coeffs, False, stream=stream)
        ms = timed(lambda: ba.barycentric.batch_eval(ba.barycentric.BASE_AT_BASE, src, n, cols, coeffs, temp, evals,
                                                     stream))
        out["barycentric_batch_eval"] = {"elems": total, "ms": round(ms, 4), "GBs": round(8 * total / ms / 1e6, 1),
                                         "hbm_frac": round(8 * total / ms / 1e6 / HBM_PEAK_GBS, 4)}
    finally:
        for d in (x, common, coeffs, evals, temp):
            d.free()
        e0.destroy()
        e1.destroy()
    return out


def run_large(ba, ctx, stream, log_n=22, cols=256, lb=2, steps=3):
    """Large trace LDE: 256 columns of 2^22 at blowup 4 (2^32 output elements, 34.4 GB), the
    whole pipeline captured as one hipGraph, evaluations kept (monomials in their own buffer)."""
    import ctypes

    lib = ba.lib()
    n, B = 1 << log_n, 1 << lb
    d_ev, d_mono, d_out = ba.DeviceAllocation(n * cols), ba.DeviceAllocation(n * cols), ba.DeviceAllocation(B * n * cols)
    try:
        ba.runtime.fill_uniform(d_ev, n, cols, 22)
        ba.check(lib.ba_stream_begin_capture(stream.handle), "begin_capture")
        ba.lde.lde_batch(ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono, stream=stream)
        graph = ctypes.c_void_p(None)
        ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(graph)), "end_capture")
        ba.check(lib.ba_graph_launch(graph, stream.handle), "graph_launch")  # warm-up
        stream.synchronize()
        e0, e1 = ba.CudaEvent(), ba.CudaEvent()
        e0.record(stream)
        for _ in range(steps):
            ba.check(lib.ba_graph_launch(graph, stream.handle), "graph_launch")
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        e0.destroy()
        e1.destroy()
        ba.check(lib.ba_graph_destroy(graph), "graph_destroy")
        # per pass kernel: one traced eager LDE (each launch covers all 256 columns, 2^30 elements)
        from boojum_amd.profile import LaunchTrace
        per = {}
        with LaunchTrace(64) as tr:
            ba.lde.lde_batch(ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono, stream=stream)
        stream.synchronize()
        for name, kms in tr.durations_ms():
            tot, cnt = per.get(name, (0.0, 0))
            per[name] = (tot + kms, cnt + 1)
        tr.destroy()
    finally:
        for t in (d_ev, d_mono, d_out):
            t.free()
    out_elems = cols * B * n
    alg = 8 * n * (1 + B) * cols
    pass_bytes = 16 * n * cols  # every pass launch reads and writes each element once
    kernels = {k: {"avg_launch_ms": round(v[0] / v[1], 3),
                   "hbm_frac": round(pass_bytes / (v[0] / v[1] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
               for k, v in per.items()}
    return {"workload": f"coset LDE, {cols} columns x 2^{log_n}, blowup {B}, one hipGraph",
            "output_bytes": out_elems * 8, "steps": steps, "ms_per_step": round(ms, 3),
            "melem_s": round(out_elems / (ms / 1e3) / 1e6, 1),
            "lde_alg_GBs": round(alg / (ms / 1e3) / 1e9, 1), "frac": round(alg / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "pass_kernels": kernels}


def run_serve(ba, ctx, n_requests=240, rate_hz=40.0, seed=5, budget_frac=0.6, slots=2, max_bypass=32):
    """Serve mode: a Poisson stream of LDE requests (8..256 columns, log_n uniform in 12..23,
    blowup 2, 4 or 8, seeded synthetic inputs generated on the device) into a server whose
    arena is 60% of HBM.  One client thread submits at the arrival times; a reaper thread
    releases each result as soon as it is done.  Latency is submit -> completion, split into
    queue time (submit -> admitted: memory granted) and the rest (admitted -> done), and
    reported per size class of output elements."""
    import threading
    import numpy as np

    free, total = ba.mem_get_info()
    budget = min(int(total * budget_frac), int(free) - (8 << 30))
    rng = np.random.default_rng(seed)
    reqs = [(int(rng.integers(12, 24)), int(rng.integers(8, 257)), int(rng.choice([1, 2, 3])))
            for _ in range(n_requests)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests))
    srv = ba.serve.Server(ctx, budget, slots)
    srv.set_max_bypass(max_bypass)
    outstanding, lock, finished = {}, threading.Lock(), threading.Event()
    results = {}

    def reaper():
        while True:
            with lock:
                pending = list(outstanding.items())
            for t, k in pending:
                r = srv.poll(t)
                if r is not None:
                    srv.release(t)
                    with lock:
                        del outstanding[t]
                        results[k] = r
            with lock:
                if finished.is_set() and not outstanding:
                    return
            time.sleep(0.0005)

    th = threading.Thread(target=reaper)
    th.start()
    t0 = time.perf_counter()
    for k, ((log_n, cols, lb), at) in enumerate(zip(reqs, arrivals)):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        t = srv.submit(log_n, cols, lb, seed=int(rng.integers(0, 2**63)))
        with lock:
            outstanding[t] = k
    finished.set()
    th.join()
    st = srv.stats()
    srv.destroy()

    def pct(v, f):
        v = sorted(v)
        return v[min(len(v) - 1, int(f * (len(v) - 1) + 0.5))] if v else None

    # size classes by output elements: small < 2^26 (512 MiB), medium < 2^30, large >= 2^30
    classes = {"small (<2^26 out)": [], "medium (2^26..2^30)": [], "large (>=2^30)": []}
    names = list(classes)
    for k, (log_n, cols, lb) in enumerate(reqs):
        r = results.get(k)
        if r is None or r["status"] != 0:
            continue
        out = cols << (log_n + lb)
        c = names[0] if out < (1 << 26) else names[1] if out < (1 << 30) else names[2]
        classes[c].append((r["latency_ms"], r["queue_ms"], r["latency_ms"] - r["queue_ms"]))
    per_class = {}
    for c, v in classes.items():
        if not v:
            continue
        lat, q, run = [x[0] for x in v], [x[1] for x in v], [x[2] for x in v]
        per_class[c] = {"requests": len(v), "p50_ms": round(pct(lat, 0.5), 3), "p99_ms": round(pct(lat, 0.99), 3),
                        "queue_p50_ms": round(pct(q, 0.5), 3), "queue_p99_ms": round(pct(q, 0.99), 3),
                        "run_p50_ms": round(pct(run, 0.5), 3), "run_p99_ms": round(pct(run, 0.99), 3)}
    allq = [r["queue_ms"] for r in results.values() if r["status"] == 0]
    allrun = [r["latency_ms"] - r["queue_ms"] for r in results.values() if r["status"] == 0]
    return {"workload": "Poisson stream of coset-LDE requests: 8..256 columns, log_n U{12..23}, blowup "
                        "2/4/8, synthetic inputs filled on the device",
            "requests": n_requests, "rate_hz": rate_hz, "slots": slots, "max_bypass": max_bypass,
            "budget_bytes": st["budget_bytes"],
            "peak_bytes": st["peak_bytes"], "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
            "mean_ms": round(st["mean_ms"], 3), "max_ms": round(st["max_ms"], 3),
            "queue_p50_ms": round(pct(allq, 0.5), 3) if allq else None,
            "queue_p99_ms": round(pct(allq, 0.99), 3) if allq else None,
            "run_p50_ms": round(pct(allrun, 0.5), 3) if allrun else None,
            "run_p99_ms": round(pct(allrun, 0.99), 3) if allrun else None,
            "per_size_class": per_class,
            "sustained_melem_s": round(st["output_elems"] / st["elapsed_s"] / 1e6, 1),
            "offered_melem_s": round(sum(c << (l + b) for l, c, b in reqs) / arrivals[-1] / 1e6, 1),
            "elapsed_s": round(st["elapsed_s"], 3), "batches": st["batches"],
            "mean_batch_jobs": round(st["mean_batch_jobs"], 2), "failed": st["failed"]}


def run_classes(ba, ctx, args):
    """Per size class log_n: a forward natural -> bit-reversed batch NTT over 2^26 elements
    (2^(26 - log_n) columns, built-in plan), args.steps launches after one warm-up; prints one
    JSON line with ms and Melem/s per class (a profiling aid for per-class counters)."""
    lo, hi = (int(x) for x in args.only_classes.split(":"))
    stream = ba.CudaStream()
    total = 1 << 26
    d_in, d_out = ba.DeviceAllocation(total), ba.DeviceAllocation(total)
    ba.runtime.fill_uniform(d_in, total, 1, 11)
    out = {}
    for log_n in range(lo, hi + 1):
        n, cols = 1 << log_n, total >> log_n
        call = lambda: ba.ntt.batch_ntt_ctx(ctx, d_in.ptr, d_out.ptr, log_n, cols, n, n, False, False,  # noqa: E731
                                            stream=stream)
        call()
        stream.synchronize()
        e0, e1 = ba.CudaEvent(), ba.CudaEvent()
        e0.record(stream)
        for _ in range(args.steps):
            call()
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / args.steps
        out[str(log_n)] = {"columns": cols, "ms": round(ms, 4), "melem_s": round(total / ms / 1e3, 1),
                           "plan": ba.tune.get_plan(ctx, log_n, False)[0]}
    d_in.free()
    d_out.free()
    stream.destroy()
    print(json.dumps({"metric": "batch NTT per size class (profiling aid)", "unit": "ms", "steps": args.steps,
                      "config": {"workload": "forward natural-order batch NTT, 2^26 elements per class"},
                      "classes": out}), flush=True)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: every box in this build has one MI355X and the LDE shards "
              "by columns with no communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.serial_cosets:
        os.environ["BOOJUM_AMD_LDE_SERIAL"] = "1"  # read by the library at its first LDE
    import boojum_amd as ba
    from boojum_amd.profile import LaunchTrace

    assert ba.device_count() >= 1, "no GPU visible"
    ba.set_device(0)
    ctx = ba.Context.create(12, 12)
    if args.only_classes:
        run_classes(ba, ctx, args)
        return
    if args.only_single:
        stream = ba.CudaStream()
        print(json.dumps({"metric": "single NTT per call", "single": run_single(ba, ctx, stream)}), flush=True)
        stream.destroy()
        return
    log_n, cols, lb = args.log_n, args.cols, args.log_blowup
    n, B = 1 << log_n, 1 << lb
    tuning = None
    if args.tune_headline:
        t0 = time.perf_counter()
        ba.tune.autotune(ctx, log_n, log_n, elems_per_trial=n * cols)
        tuning = {"log_n_range": [log_n, log_n], "seconds": round(time.perf_counter() - t0, 2),
                  f"plan_{log_n}_natural": ba.tune.get_plan(ctx, log_n, False),
                  f"plan_{log_n}_bitrev": ba.tune.get_plan(ctx, log_n, True)}
    d_ev = ba.DeviceAllocation(n * cols)
    d_mono = ba.DeviceAllocation(n * cols)
    d_out = ba.DeviceAllocation(B * n * cols)
    ba.runtime.fill_uniform(d_ev, n, cols, 1)
    stream = ba.CudaStream()
    if args.calibrate_copy:
        cn = 1 << 27
        c_src, c_dst = ba.DeviceAllocation(cn), ba.DeviceAllocation(cn)
        ba.runtime.fill_uniform(c_src, cn, 1, 3)
        for _ in range(3):
            ba.runtime.field_op(3, c_src, None, c_dst, cn)
        ba.synchronize()
        c_src.free()
        c_dst.free()
        butterfly_rate(ba, stream)  # so the counter passes also cover the register-only butterfly loop

    def step_eager(s=stream):
        ba.lde.lde_batch(ctx, d_ev, log_n, cols, lb, d_out, monomials=d_mono, stream=s)

    # capture the whole LDE (iNTT, coset scale, forward NTTs) as one graph
    import ctypes
    lib = ba.lib()
    ba.check(lib.ba_stream_begin_capture(stream.handle), "begin_capture")
    step_eager()
    graph = ctypes.c_void_p(None)
    ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(graph)), "end_capture")

    def step():
        ba.check(lib.ba_graph_launch(graph, stream.handle), "graph_launch")

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    ba.synchronize()
    e0, e1 = ba.CudaEvent(), ba.CudaEvent()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    e1.synchronize()
    ba.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_ms(e1) / args.steps
    out_elems = cols * B * n
    value = out_elems / (ms_per_step / 1e3) / 1e6
    lde_alg_bytes = 8 * n * (1 + B) * cols

    # dominant kernel: one traced eager step, per-launch HIP events on the launch stream
    per_kernel = {}
    trace_steps = max(1, min(5, args.steps))
    with LaunchTrace(64 * trace_steps) as tr:
        for _ in range(trace_steps):
            step_eager()
    stream.synchronize()
    for name, ms in tr.durations_ms():
        tot, cnt = per_kernel.get(name, (0.0, 0))
        per_kernel[name] = (tot + ms, cnt + 1)
    tr.destroy()
    dom_name, (dom_tot, dom_cnt) = max(per_kernel.items(), key=lambda kv: kv[1][0])
    dom_avg_ms = dom_tot / dom_cnt
    # each pass launch reads and writes every element of the batch once: 16 B per element
    dom_bytes = 16 * n * cols
    # the stages (index bits) the dominant pass transforms: its S template argument
    import re
    m = re.search(r"ntt_pass_(?:contig|strided)<(\d+),", dom_name)
    dom_stages = int(m.group(1)) if m else 0
    dom_bfly = dom_stages * n * cols // 2
    achieved = dom_bytes / (dom_avg_ms / 1e3) / 1e9
    traffic = None
    valu = None
    prof_source = None
    if (log_n, cols, lb) == (20, 64, 1) and os.path.exists(TRAFFIC_FILE):
        prof = json.load(open(TRAFFIC_FILE))
        t = prof["traffic_bytes_per_launch"].get(dom_name)
        traffic = int(t) if t is not None else None
        # the counters are not measurable inside this run: they come from the committed rocprofv3
        # summary, tagged with the commit of the kernels it measured
        prof_source = {"file": os.path.relpath(TRAFFIC_FILE, ROOT), "measured_commit": prof.get("measured_commit"),
                       "counters": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE / SQ_INSTS_VALU"}
        v = prof.get("valu", {}).get(dom_name)
        bf = prof.get("butterfly_rate")
        if v is not None and bf is not None:
            # integer VALU work of the same kernel: instructions per element (SQ_INSTS_VALU, one
            # wave instruction covers 64 lanes) over this run's launch time, against the rate the
            # same arithmetic reaches in registers, measured now: ba_diag_butterfly_rate's
            # butterflies/s times its instructions per butterfly (SQ_INSTS_VALU, committed profile)
            winst = v["valu_insts_per_elem"] * n * cols / 64
            achieved_w = winst / (dom_avg_ms / 1e3)
            bfly_s, _ = butterfly_rate(ba, stream)
            ceiling = bfly_s * bf["valu_insts_per_butterfly"] / 64
            valu = {"bound": "valu", "insts_per_elem": v["valu_insts_per_elem"],
                    "achieved": round(achieved_w / 1e9, 1), "peak": round(ceiling / 1e9, 1),
                    "unit": "G wave-instr/s", "frac": round(achieved_w / ceiling, 4),
                    "peak_kind": "measured: the passes' butterfly arithmetic in registers "
                                 f"({bf['valu_insts_per_butterfly']} VALU instructions per butterfly, "
                                 f"{bfly_s / 1e9:.0f} G butterflies/s)",
                    "model_peak": VALU_PEAK_WINST / 1e9,
                    "model_frac": round(achieved_w / VALU_PEAK_WINST, 4)}

    # the same graph with the contiguous pass pinned at 3 and at 4 waves per SIMD (the built-in
    # choice above takes 4 at this size); steps timed as above
    contig_waves = {}
    if not args.tune_headline and (log_n, cols, lb) == (20, 64, 1):
        for waves in (3, 4):
            for br in (False, True):
                bits, _ = ba.tune.get_plan(ctx, log_n, br)
                ba.tune.set_plan(ctx, log_n, br, bits, None, contig_waves=waves)
            ba.check(lib.ba_stream_begin_capture(stream.handle), "begin_capture")
            step_eager()
            g2 = ctypes.c_void_p(None)
            ba.check(lib.ba_stream_end_capture(stream.handle, ctypes.byref(g2)), "end_capture")
            for _ in range(10):
                ba.check(lib.ba_graph_launch(g2, stream.handle), "graph_launch")
            v0, v1 = ba.CudaEvent(), ba.CudaEvent()
            v0.record(stream)
            for _ in range(args.steps):
                ba.check(lib.ba_graph_launch(g2, stream.handle), "graph_launch")
            v1.record(stream)
            v1.synchronize()
            contig_waves[str(waves)] = round(v0.elapsed_ms(v1) / args.steps, 4)
            ba.check(lib.ba_graph_destroy(g2), "graph_destroy")
        for br in (False, True):
            ba.tune.set_plan(ctx, log_n, br, None)

    per_call = None
    if not args.no_per_call:
        # one ba_lde_batch call per column (6 launches each), eager, as a caller without batching
        pc_steps = max(1, min(3, args.steps))
        for _ in range(1):
            for k in range(cols):
                ba.lde.lde_batch(ctx, d_ev.sub(k * n, n), log_n, 1, lb, d_out.sub(k * B * n, B * n),
                                 monomials=d_mono.sub(k * n, n), stream=stream)
        stream.synchronize()
        p0, p1 = ba.CudaEvent(), ba.CudaEvent()
        p0.record(stream)
        for _ in range(pc_steps):
            for k in range(cols):
                ba.lde.lde_batch(ctx, d_ev.sub(k * n, n), log_n, 1, lb, d_out.sub(k * B * n, B * n),
                                 monomials=d_mono.sub(k * n, n), stream=stream)
        p1.record(stream)
        p1.synchronize()
        pc_ms = p0.elapsed_ms(p1) / pc_steps
        per_call = {"ms_per_step": round(pc_ms, 4), "melem_s": round(out_elems / (pc_ms / 1e3) / 1e6, 1),
                    "batched_speedup": round(pc_ms / ms_per_step, 3)}

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks field, exact)",
        "data": "synthetic: uniform field elements from a seeded SplitMix64 stream, generated in HBM",
        "config": {"workload": f"batched coset LDE, {cols} columns x 2^{log_n}, blowup {B}, shift 7, "
                               f"bit-reversed output, one hipGraph per step",
                   "columns": cols, "log_n": log_n, "blowup": B, "parallelism": "single GPU",
                   "cosets": "serial, one stream" if args.serial_cosets else "concurrent, auxiliary streams"},
        "lde_roofline": {"alg_bytes_per_step": lde_alg_bytes,
                         "achieved_GBs": round(lde_alg_bytes / (ms_per_step / 1e3) / 1e9, 1),
                         "frac": round(lde_alg_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_launch_ms": round(dom_avg_ms, 4),
                     "alg_bytes_per_launch": dom_bytes,
                     # which ceiling the kernel is nearer: its VALU rate against the measured VALU
                     # ceiling, or its HBM rate against the HBM peak (None without the VALU counters)
                     "limited_by": (None if valu is None else
                                    "valu" if valu["frac"] > achieved / HBM_PEAK_GBS else "hbm"),
                     # useful work: radix-2 butterflies (n/2 per stage per column) per second
                     "butterflies_per_s": round(dom_bfly / (dom_avg_ms / 1e3), 1),
                     "element_stages_per_s": round(2 * dom_bfly / (dom_avg_ms / 1e3), 1),
                     "counters_source": prof_source},
        "valu_roofline": valu,
        "ms_per_step_contig_waves": contig_waves or None,
        "per_kernel_ms": {k: round(v[0] / v[1], 4) for k, v in per_kernel.items()},
        "per_call": per_call,
        "tuning": tuning,
        "wall_s": round(wall, 3),
    }
    if not args.no_uneven:
        # its own context, autotuned over the batch's size classes (tuning launches run before
        # and apart from every timed or traced region)
        uctx = ba.Context.create(12, 12)
        utune = None
        if not args.no_tune:
            t0 = time.perf_counter()
            ba.tune.autotune(uctx, 12, 22, elems_per_trial=1 << 26)
            utune = {"log_n_range": [12, 22], "seconds": round(time.perf_counter() - t0, 2),
                     # per size class and input order: pass split (low bits first), contiguous-pass
                     # waves per SIMD, tile (10: 1024-element latency-shaped tiles at every batch
                     # size, 0: by batch size)
                     # (n: natural-input NTTs, b: bit-reversed input, l: the LDE's coset NTTs)
                     "choices": {f"{ln}{'nbl'[br]}": [ba.tune.get_plan(uctx, ln, br)[0],
                                                     ba.tune.get_contig_waves(uctx, ln, br),
                                                     ba.tune.get_tile_log(uctx, ln, br)]
                                 for ln in range(12, 23) for br in (0, 1, ba.tune.LDE_COSETS)}}
        line["uneven"] = run_uneven(ba, uctx, stream)
        line["uneven"]["tuning"] = utune
        uctx.destroy()
    if not args.no_single:
        line["single"] = run_single(ba, ctx, stream)
    if not args.no_large:
        line["large"] = run_large(ba, ctx, stream)
    if not args.no_harness:
        line["harness"] = run_harness(ba, stream)
    if not args.no_merkle:
        line["merkle"] = run_merkle(ba, ctx, stream, d_out, log_n, cols, lb)
    if not args.no_ops:
        line["ops"] = run_ops(ba, stream, d_ev, d_mono, log_n, cols)
    if not args.no_serve:
        # the default policy (later requests that fit pass a head waiting for memory, at most 32 per
        # head, two slots), then strict arrival order and four slots on the same request stream
        # on the server's own context, autotuned over the stream's size classes before the first
        # request (as a server would at start-up; tuning runs apart from every timed region)
        sctx = ba.Context.create(12, 12)
        stune = None
        if not args.no_tune:
            t0 = time.perf_counter()
            ba.tune.autotune(sctx, 12, 23, elems_per_trial=1 << 26)
            stune = {"log_n_range": [12, 23], "seconds": round(time.perf_counter() - t0, 2)}
        line["serve"] = run_serve(ba, sctx)
        line["serve"]["tuning"] = stune
        line["serve"]["variants"] = {
            "strict_order_2_slots": run_serve(ba, sctx, max_bypass=0),
            "bypass_4_slots": run_serve(ba, sctx, slots=4)}
        sctx.destroy()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(log_n, lb)
    ba.check(lib.ba_graph_destroy(graph), "graph_destroy")
    for t in (d_ev, d_mono, d_out):
        t.free()
    stream.destroy()
    ctx.destroy()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
