"""Autotuning of pass splits, columns per workgroup and the contiguous pass's waves per SIMD
(include/boojum_amd.h: ba_autotune,
ba_tuning_get, ba_tuning_set).  No reference counterpart: boojum-cuda's kernel shapes are fixed
tables (crates/boojum-cuda/src/ntt.rs:57-176)."""
import ctypes

from ._lib import TunedPlan, check, lib

# bitrev_inputs value that addresses the LDE's coset NTTs' plan (BA_TUNING_LDE_COSETS)
LDE_COSETS = 2


def autotune(ctx, log_n_min, log_n_max, elems_per_trial=1 << 26, stream=None):
    """Times every valid split and per-pass reps for each size class in the range; the fastest
    becomes the plan every later transform on `ctx` runs with."""
    check(lib().ba_autotune(ctx.handle, log_n_min, log_n_max, elems_per_trial,
                            stream.handle if stream is not None else None), "autotune")


def get_plan(ctx, log_n, bitrev_inputs=False):
    """(bits, reps): the pass split from the contiguous low pass upwards, and the columns per
    workgroup of each pass (0 = built-in choice)."""
    p = TunedPlan()
    check(lib().ba_tuning_get(ctx.handle, log_n, int(bitrev_inputs), ctypes.byref(p)), "tuning_get")
    n = p.num_passes
    return list(p.bits[:n]), list(p.reps[:n])


def get_contig_waves(ctx, log_n, bitrev_inputs=False):
    """Waves per SIMD of the size class's 9..12-bit contiguous pass (3, 4; 0 = built-in choice)."""
    p = TunedPlan()
    check(lib().ba_tuning_get(ctx.handle, log_n, int(bitrev_inputs), ctypes.byref(p)), "tuning_get")
    return p.contig_waves


def get_tile_log(ctx, log_n, bitrev_inputs=False):
    """Tile shape of the size class's plain batch NTTs: 10 (latency-shaped 1024-element tiles at every
    batch size), 12 (4096-element tiles at every batch size) or 0 (built-in: by batch size)."""
    p = TunedPlan()
    check(lib().ba_tuning_get(ctx.handle, log_n, int(bitrev_inputs), ctypes.byref(p)), "tuning_get")
    return p.tile_log


def set_plan(ctx, log_n, bitrev_inputs, bits, reps=None, contig_waves=0, tile_log=0):
    """Sets the plan of a size class; bits=None restores the built-in plan."""
    if bits is None:
        check(lib().ba_tuning_set(ctx.handle, log_n, int(bitrev_inputs), None), "tuning_set")
        return
    p = TunedPlan()
    p.num_passes = len(bits)
    for i, b in enumerate(bits[:8]):
        p.bits[i] = b
    for i, r in enumerate((reps or [])[:8]):
        p.reps[i] = r
    p.contig_waves = contig_waves
    p.tile_log = tile_log
    check(lib().ba_tuning_set(ctx.handle, log_n, int(bitrev_inputs), ctypes.byref(p)), "tuning_set")


def library_candidates(log_n):
    """The splits ba_autotune tries, as the library enumerates them (host call, no device)."""
    n = ctypes.c_uint32(0)
    check(lib().ba_tuning_candidates(log_n, None, 0, ctypes.byref(n)), "tuning_candidates")
    arr = (TunedPlan * max(1, n.value))()
    check(lib().ba_tuning_candidates(log_n, arr, n.value, ctypes.byref(n)), "tuning_candidates")
    return [list(arr[i].bits[:arr[i].num_passes]) for i in range(n.value)]


def candidate_splits(log_n):
    """The splits ba_autotune tries (same enumeration as csrc/tune.hip), low pass first."""
    if log_n <= 12:
        return [[log_n]]
    out = []

    def comps(rest, parts, cur):
        if parts == 0:
            if rest == 0:
                out.append(list(cur))
            return
        for b in range(4, min(10, rest) + 1):
            comps(rest - b, parts - 1, cur + [b])

    for sc in range(9, 13):
        if log_n < sc + 4:
            continue
        rest = log_n - sc
        min_parts = (rest + 9) // 10
        for parts in range(min_parts, min(min_parts + 1, 7) + 1):
            before = len(out)
            comps(rest, parts, [])
            for k in range(before, len(out)):
                out[k] = [sc] + out[k]
    return out
