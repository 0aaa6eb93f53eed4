"""ctypes view of the CPU oracle (oracle/libgl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, where it is the checker or the timed CPU baseline.  Nothing on the
product path imports it.

The C restatement follows boojum's CPU FFT as the reference test calls it
(crates/boojum-cuda/src/ntt.rs:391-611) and shivini's LDE composition
(crates/shivini/src/data_structures/storage.rs:154-260, primitives/ntt.rs:10-42).
"""
import ctypes
import os

import numpy as np

P = 0xFFFFFFFF00000001
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgl_oracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(_LIB_PATH)
        u64, u32, i32 = ctypes.c_uint64, ctypes.c_uint, ctypes.c_int
        for name, res, args in [
            ("gl_reduce", u64, [u64]),
            ("gl_add", u64, [u64, u64]),
            ("gl_sub", u64, [u64, u64]),
            ("gl_mul", u64, [u64, u64]),
            ("gl_pow", u64, [u64, u64]),
            ("gl_inv", u64, [u64]),
            ("gl_omega", u64, [u32]),
            ("gl_coset_shift", u64, [u32, u32, u32]),
            ("gl_precompute_twiddles", None, [_u64p, u32, i32]),
            ("gl_oracle_batch_ntt", i32, [_u64p, _u64p, u32, u32, u64, u64, i32, i32, u32, u32, i32]),
            ("gl_oracle_bit_reverse", None, [_u64p, u64, _u64p, u64, u32, u32]),
            ("gl_oracle_lde", i32, [_u64p, u64, _u64p, u64, u64, u32, u32, u32, i32]),
            ("gl_fill_uniform", None, [_u64p, ctypes.c_size_t, u64]),
            ("gl_oracle_powers_of_w", None, [u32, u32, i32, i32, _u64p, u32]),
            ("gl_oracle_powers_of_g", None, [u32, u32, i32, i32, _u64p, u32]),
            ("gl_oracle_powers_by_val_bf", None, [u64, u32, i32, _u64p, u32]),
            ("gl_oracle_powers_by_val_ef", None, [u64, u64, u32, i32, _u64p, u32]),
            ("gl_oracle_omega_shift", None, [_u64p, u32, u32, i32, u32, _u64p, u32]),
            ("gl_oracle_barycentric_coeffs", None, [i32, u64, u64, u64, u32, i32, _u64p]),
            ("gl_oracle_barycentric_eval", None, [i32, _u64p, u32, u32, _u64p, _u64p]),
            ("gl_oracle_poseidon2_permutation", None, [_u64p, _u64p]),
            ("gl_oracle_poseidon2_matrix", None, [i32, _u64p]),
            ("gl_oracle_poseidon2_leaves", None, [_u64p, _u64p, u32, u32, u32, i32, i32, _u64p]),
            ("gl_oracle_poseidon2_nodes", None, [_u64p, _u64p, u32, _u64p]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def field_op(name, *args):
    return int(getattr(lib(), name)(*args))


def omega(log_size):
    return int(lib().gl_omega(log_size))


def coset_shift(log_n, log_ext, coset_idx):
    return int(lib().gl_coset_shift(log_n, log_ext, coset_idx))


def twiddles(log_n, inverse):
    tw = np.zeros(max(1, (1 << log_n) // 2), dtype=np.uint64)
    lib().gl_precompute_twiddles(_ptr(tw), log_n, int(inverse))
    return tw[: (1 << log_n) // 2]


def batch_ntt(inputs, log_n, num_ntts, stride_in, stride_out, bitrev_inputs, inverse,
              log_ext=0, coset_idx=0, out=None, nthreads=1):
    """Out-of-place (or in place when out is inputs) batch NTT with the semantics of
    batch_ntt_internal (crates/boojum-cuda/src/ntt.rs:179-291)."""
    inputs = np.ascontiguousarray(inputs, dtype=np.uint64)
    if out is None:
        out = np.zeros(max(num_ntts * stride_out, 1), dtype=np.uint64)
    rc = lib().gl_oracle_batch_ntt(_ptr(inputs), _ptr(out), log_n, num_ntts, stride_in, stride_out,
                                   int(bitrev_inputs), int(inverse), log_ext, coset_idx, nthreads)
    if rc != 0:
        raise ValueError("invalid batch_ntt arguments")
    return out


def bit_reverse(src, log_rows, cols, src_stride=None, dst_stride=None, out=None):
    n = 1 << log_rows
    src_stride = n if src_stride is None else src_stride
    dst_stride = n if dst_stride is None else dst_stride
    src = np.ascontiguousarray(src, dtype=np.uint64)
    if out is None:
        out = np.zeros(cols * dst_stride, dtype=np.uint64)
    lib().gl_oracle_bit_reverse(_ptr(src), src_stride, _ptr(out), dst_stride, log_rows, cols)
    return out


def lde(evals, log_n, num_cols, log_blowup, in_stride=None, out_col_stride=None,
        out_coset_stride=None, out=None, nthreads=1):
    n = 1 << log_n
    in_stride = n if in_stride is None else in_stride
    out_col_stride = (n << log_blowup) if out_col_stride is None else out_col_stride
    out_coset_stride = n if out_coset_stride is None else out_coset_stride
    evals = np.ascontiguousarray(evals, dtype=np.uint64)
    if out is None:
        size = (num_cols - 1) * out_col_stride + ((1 << log_blowup) - 1) * out_coset_stride + n
        out = np.zeros(size, dtype=np.uint64)
    rc = lib().gl_oracle_lde(_ptr(evals), in_stride, _ptr(out), out_col_stride, out_coset_stride,
                             log_n, num_cols, log_blowup, nthreads)
    if rc != 0:
        raise ValueError("invalid lde arguments")
    return out


def fill_uniform(n, seed):
    v = np.empty(n, dtype=np.uint64)
    lib().gl_fill_uniform(_ptr(v), n, seed)
    return v


def powers_of_w(log_degree, offset, inverse, bit_reverse, count):
    r = np.empty(count, dtype=np.uint64)
    lib().gl_oracle_powers_of_w(log_degree, offset, int(inverse), int(bit_reverse), _ptr(r), count)
    return r


def powers_of_g(log_degree, offset, inverse, bit_reverse, count):
    r = np.empty(count, dtype=np.uint64)
    lib().gl_oracle_powers_of_g(log_degree, offset, int(inverse), int(bit_reverse), _ptr(r), count)
    return r


def powers_by_val_bf(base, offset, bit_reverse, count):
    r = np.empty(count, dtype=np.uint64)
    lib().gl_oracle_powers_by_val_bf(base, offset, int(bit_reverse), _ptr(r), count)
    return r


def powers_by_val_ef(base_c0, base_c1, offset, bit_reverse, count):
    """Two planes: [c0 of every power, c1 of every power]."""
    r = np.empty(2 * count, dtype=np.uint64)
    lib().gl_oracle_powers_by_val_ef(base_c0, base_c1, offset, int(bit_reverse), _ptr(r), count)
    return r


def omega_shift(src, log_degree, offset, inverse, shift):
    src = np.ascontiguousarray(src, dtype=np.uint64)
    dst = np.empty_like(src)
    lib().gl_oracle_omega_shift(_ptr(src), log_degree, offset, int(inverse), shift, _ptr(dst), src.size)
    return dst


def barycentric_coeffs(x, coset, count, bit_reverse, ext=False):
    """x: an int (base field) or a (c0, c1) pair (extension)."""
    x0, x1 = (x if ext else (x, 0))
    r = np.empty(count * (2 if ext else 1), dtype=np.uint64)
    lib().gl_oracle_barycentric_coeffs(int(ext), x0, x1, coset, count, int(bit_reverse), _ptr(r))
    return r


def barycentric_eval(kind, ys, count, num_polys, coeffs):
    ys = np.ascontiguousarray(ys, dtype=np.uint64)
    coeffs = np.ascontiguousarray(coeffs, dtype=np.uint64)
    r = np.empty(num_polys * (1 if kind == 0 else 2), dtype=np.uint64)
    lib().gl_oracle_barycentric_eval(kind, _ptr(ys), count, num_polys, _ptr(coeffs), _ptr(r))
    return r


# ---- Poseidon2 (width 12) Merkle hashing; round_constants: 360 u64 (30 rounds x 12 lanes) ----
def poseidon2_permutation(state, round_constants):
    st = np.ascontiguousarray(np.array(state, dtype=np.uint64)).copy()
    assert st.size == 12
    lib().gl_oracle_poseidon2_permutation(_ptr(st), _ptr(np.ascontiguousarray(round_constants, dtype=np.uint64)))
    return st


def poseidon2_matrix(internal, x):
    v = np.ascontiguousarray(np.array(x, dtype=np.uint64)).copy()
    lib().gl_oracle_poseidon2_matrix(int(internal), _ptr(v))
    return v


def poseidon2_leaves(values, rows_count, cols_count, count, round_constants, load_intermediate=False,
                     store_intermediate=False, results=None):
    values = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros(4 * count, dtype=np.uint64) if results is None else np.ascontiguousarray(results, dtype=np.uint64).copy()
    lib().gl_oracle_poseidon2_leaves(_ptr(values), _ptr(out), rows_count, cols_count, count, int(load_intermediate),
                                     int(store_intermediate), _ptr(np.ascontiguousarray(round_constants, dtype=np.uint64)))
    return out


def poseidon2_nodes(values, count, round_constants):
    values = np.ascontiguousarray(values, dtype=np.uint64)
    out = np.zeros(4 * count, dtype=np.uint64)
    lib().gl_oracle_poseidon2_nodes(_ptr(values), _ptr(out), count,
                                    _ptr(np.ascontiguousarray(round_constants, dtype=np.uint64)))
    return out


def poseidon2_merkle_tree(values, log_rows_per_hash, leaves_count, layers_count, round_constants):
    """build_merkle_tree's layout (crates/boojum-cuda/src/poseidon2.rs:253-316): 8 x leaves_count
    words; the leaf digests (4 planes of leaves_count) first, then each node layer (4 planes of
    half the previous count) after the previous one; the words past the last layer are zero."""
    out = np.zeros(8 * leaves_count, dtype=np.uint64)
    rows = 1 << log_rows_per_hash
    cols = values.size // (leaves_count * rows)
    layer = poseidon2_leaves(values, rows, cols, leaves_count, round_constants)
    out[:layer.size] = layer
    at, count = layer.size, leaves_count
    for _ in range(layers_count - 1):
        count //= 2
        layer = poseidon2_nodes(layer, count, round_constants)
        out[at:at + layer.size] = layer
        at += layer.size
    return out
