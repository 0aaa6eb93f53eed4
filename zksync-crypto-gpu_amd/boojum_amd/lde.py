"""Batched coset LDE (include/boojum_amd.h: ba_lde_batch), the composition shivini drives
through boojum_cuda::ntt (crates/shivini/src/data_structures/storage.rs:154-260)."""
from ._lib import check, lib


NATURAL_MONOMIALS = 1  # BA_LDE_NATURAL_MONOMIALS


def lde_batch(ctx, evals, log_n, num_cols, log_blowup, out, in_stride=None, monomials=None,
              out_col_stride=None, out_coset_stride=None, stream=None, natural_monomials=False):
    n = 1 << log_n
    args = (ctx.handle, evals.ptr, n if in_stride is None else in_stride,
            monomials.ptr if monomials is not None else None, out.ptr,
            (n << log_blowup) if out_col_stride is None else out_col_stride,
            n if out_coset_stride is None else out_coset_stride, log_n, num_cols, log_blowup)
    h = stream.handle if stream is not None else None
    if natural_monomials:
        check(lib().ba_lde_batch_ex(*args, NATURAL_MONOMIALS, h), "lde_batch_ex")
    else:
        check(lib().ba_lde_batch(*args, h), "lde_batch")
