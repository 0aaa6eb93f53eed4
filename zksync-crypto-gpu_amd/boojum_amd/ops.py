"""boojum_cuda::ops_complex::bit_reverse mirror (crates/boojum-cuda/src/ops_complex.rs:262-331)."""
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
