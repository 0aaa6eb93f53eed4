"""boojum_cuda::ops_complex mirror: bit_reverse (crates/boojum-cuda/src/ops_complex.rs:262-331),
get_powers_of_w/g, get_powers_by_val/ref and omega_shift (ops_complex.rs:31-260).  DeviceSlices
carry (pointer, length); `result.len` is the count, as the Rust functions take it from the slice;
extension-field results hold 2 * count elements (two planes)."""
from ._lib import check, lib


def _h(stream):
    return stream.handle if stream is not None else None


def bit_reverse(src, dst, log_rows, cols, src_stride=None, dst_stride=None, stream=None):
    rows = 1 << log_rows
    check(lib().ba_bit_reverse(src.ptr, rows if src_stride is None else src_stride, dst.ptr,
                               rows if dst_stride is None else dst_stride, log_rows, cols,
                               _h(stream)), "bit_reverse")


def bit_reverse_in_place(values, log_rows, cols, stride=None, stream=None):
    rows = 1 << log_rows
    check(lib().ba_bit_reverse_in_place(values.ptr, rows if stride is None else stride, log_rows,
                                        cols, _h(stream)), "bit_reverse_in_place")


def get_powers_of_w(log_degree, offset, inverse, bit_reverse, result, stream=None):
    check(lib().ba_get_powers_of_w(log_degree, offset, int(inverse), int(bit_reverse), result.ptr, result.len,
                                   _h(stream)), "get_powers_of_w")


def get_powers_of_g(log_degree, offset, inverse, bit_reverse, result, stream=None):
    check(lib().ba_get_powers_of_g(log_degree, offset, int(inverse), int(bit_reverse), result.ptr, result.len,
                                   _h(stream)), "get_powers_of_g")


def get_powers_by_val_bf(base, offset, bit_reverse, result, stream=None):
    check(lib().ba_get_powers_by_val_bf(base, offset, int(bit_reverse), result.ptr, result.len, _h(stream)),
          "get_powers_by_val")


def get_powers_by_ref_bf(base, offset, bit_reverse, result, stream=None):
    """base: a DeviceSlice holding one element (a DeviceVariable in the reference)."""
    check(lib().ba_get_powers_by_ref_bf(base.ptr, offset, int(bit_reverse), result.ptr, result.len, _h(stream)),
          "get_powers_by_ref")


def get_powers_by_val_ef(base, offset, bit_reverse, result, stream=None):
    """base: (c0, c1); result: 2 * count elements."""
    check(lib().ba_get_powers_by_val_ef(base[0], base[1], offset, int(bit_reverse), result.ptr, result.len // 2,
                                        _h(stream)), "get_powers_by_val")


def get_powers_by_ref_ef(base, offset, bit_reverse, result, stream=None):
    check(lib().ba_get_powers_by_ref_ef(base.ptr, offset, int(bit_reverse), result.ptr, result.len // 2,
                                        _h(stream)), "get_powers_by_ref")


def omega_shift(src, log_degree, offset, inverse, shift, dst, stream=None):
    assert src.len == dst.len  # ops_complex.rs:237
    check(lib().ba_omega_shift(src.ptr, log_degree, offset, int(inverse), shift, dst.ptr, dst.len, _h(stream)),
          "omega_shift")


def omega_shift_in_place(values, log_degree, offset, inverse, shift, stream=None):
    check(lib().ba_omega_shift_in_place(values.ptr, log_degree, offset, int(inverse), shift, values.len,
                                        _h(stream)), "omega_shift_in_place")
