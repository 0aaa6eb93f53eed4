"""Writes tests/golden/ntt_golden.npz: seeded inputs and expected outputs for the batch NTT
(all four n2b/b2n x forward/inverse modes, with and without coset LDE factors) and for the
shivini-style LDE.

Expected outputs come from the C oracle (oracle/libgl_oracle.so) and every one of them is
re-derived here from the transform definitions with Python big integers
(tests/bigint_ref.py) before the file is written; the script refuses to write a fixture the
two disagree on.  The reference holds no NTT golden vectors of its own (its tests compare
against boojum at run time with unseeded inputs, crates/boojum-cuda/src/ntt.rs:417).

Run from the repository root:  python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import bigint_ref as br  # noqa: E402
import gl_oracle  # noqa: E402

NTT_LOG_N = range(1, 10)
LDE_LOG_N = range(1, 8)
COSET_CASES = [(0, 0), (3, 3), (1, 0)]  # (log_extension_degree, coset_index)
MODES = [(False, False), (True, False), (False, True), (True, True)]  # (bitrev_inputs, inverse)
LDE_BLOWUPS = [1, 2, 3]


def raw_inputs(n, seed):
    """Raw u64 (the reference test feeds unreduced rng output, ntt.rs:424) with the
    non-canonical edge values p, p+1 and 2^64-1 planted."""
    v = np.random.default_rng(seed).integers(0, 2**64, size=n, dtype=np.uint64)
    edges = [br.P, br.P + 1, 2**64 - 1, 0]
    for i, e in enumerate(edges[: max(0, n - 1)]):
        v[(i * 7 + 1) % n] = np.uint64(e)
    return v


def main():
    out = {}
    for log_n in NTT_LOG_N:
        n = 1 << log_n
        x = raw_inputs(n, 1000 + log_n)
        out[f"in_{log_n}"] = x
        xs = [int(t) for t in x]
        for bitrev_inputs, inverse in MODES:
            for log_ext, coset_idx in COSET_CASES:
                y = gl_oracle.batch_ntt(x, log_n, 1, n, n, bitrev_inputs, inverse, log_ext, coset_idx)
                ref = br.batch_mode(xs, log_n, bitrev_inputs, inverse, log_ext, coset_idx)
                assert [int(t) for t in y] == ref, (log_n, bitrev_inputs, inverse, log_ext, coset_idx)
                key = f"ntt_{log_n}_{int(bitrev_inputs)}{int(inverse)}_{log_ext}_{coset_idx}"
                out[key] = y
    for log_n in LDE_LOG_N:
        n = 1 << log_n
        ev = raw_inputs(n, 2000 + log_n)
        out[f"lde_in_{log_n}"] = ev
        coeffs = br.dft([int(t) for t in ev], log_n, 1, inverse=True)
        for lb in LDE_BLOWUPS:
            y = gl_oracle.lde(ev, log_n, 1, lb)
            ref = [br.horner(coeffs, br.lde_point(log_n, lb, q)) for q in range(n << lb)]
            assert [int(t) for t in y] == ref, (log_n, lb)
            out[f"lde_{log_n}_{lb}"] = y
    path = os.path.join(ROOT, "tests", "golden", "ntt_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
