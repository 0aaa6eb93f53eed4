"""boojum_cuda::barycentric mirror (crates/boojum-cuda/src/barycentric.rs:83-300) over the C-ABI:
precompute_lagrange_coeffs (at a base-field or extension point) and batch_eval (base at base,
base at extension, extension at extension) with caller-provided temp storage."""
import ctypes

from ._lib import check, lib

BASE_AT_BASE, BASE_AT_EXT, EXT_AT_EXT = 0, 1, 2


def _h(stream):
    return stream.handle if stream is not None else None


def precompute_lagrange_coeffs(x, common_factor_storage, coset, lagrange_coeffs, bit_reverse, ext=False, stream=None):
    """x, common_factor_storage: DeviceSlices of 1 (base) or 2 (extension) elements;
    lagrange_coeffs: count (base) or 2 * count (extension planes) elements."""
    count = lagrange_coeffs.len // (2 if ext else 1)
    check(lib().ba_barycentric_precompute_lagrange_coeffs(int(ext), x.ptr, common_factor_storage.ptr, coset,
                                                          lagrange_coeffs.ptr, count, int(bit_reverse), _h(stream)),
          "precompute_lagrange_coeffs")


def batch_eval_temp_elems(kind, count, num_polys):
    n = ctypes.c_uint64(0)
    check(lib().ba_barycentric_batch_eval_temp_elems(kind, count, num_polys, ctypes.byref(n)), "batch_eval_temp")
    return n.value


def batch_eval(kind, ys, count, num_polys, lagrange_coeffs, temp, evals, stream=None):
    check(lib().ba_barycentric_batch_eval(kind, ys.ptr, count, num_polys, lagrange_coeffs.ptr, temp.ptr, temp.len,
                                          evals.ptr, _h(stream)), "batch_eval")


NOALLOC_TEMP_ELEMS = 1 << 19  # the library's per-GPU array behind batch_eval_noalloc


def batch_eval_noalloc(kind, ys, count, num_polys, lagrange_coeffs, evals, stream=None):
    """batch_eval with the partial sums in the library's per-GPU array: no temp argument, no
    allocation; calls that overlap on one GPU race."""
    check(lib().ba_barycentric_batch_eval_noalloc(kind, ys.ptr, count, num_polys, lagrange_coeffs.ptr, evals.ptr,
                                                  _h(stream)), "batch_eval_noalloc")
