"""Helpers shared by the GPU parity tests."""
import sys

import numpy as np

P = 0xFFFFFFFF00000001
COLUMN_STEP = 0xD1B54A32D192ED03


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def raw_u64(rng, n):
    """Unreduced u64 inputs, as the reference test feeds (crates/boojum-cuda/src/ntt.rs:424)."""
    return rng.integers(0, 2**64, size=n, dtype=np.uint64)


def column_seed(seed, k):
    return (seed + k * COLUMN_STEP) % 2**64


def bitrev_perm(log_n):
    n = 1 << log_n
    idx = np.arange(n, dtype=np.uint64)
    r = np.zeros(n, dtype=np.uint64)
    for b in range(log_n):
        r |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(log_n - 1 - b)
    return r.astype(np.int64)
