//! `boojum_cuda::ntt` (crates/boojum-cuda/src/ntt.rs:179-369): same names, argument lists and
//! types (generic over the element representation); plus the LDE shivini composes from it
//! (crates/shivini/src/data_structures/storage.rs:154-260).
use crate::context::Context;
use crate::cudart::result::CudaResult;
use crate::cudart::slice::DeviceSlice;
use crate::cudart::stream::CudaStream;
use crate::{check, raw, raw_mut, sys, GoldilocksRepr};

pub use crate::context::OMEGA_LOG_ORDER;

/// `inputs_ptr_in` / `outputs_ptr` are device pointers covering `num_ntts` strided arrays of
/// 2^log_n elements (the reference's function is safe with raw pointers too, ntt.rs:179-191).
#[allow(clippy::too_many_arguments)]
pub fn batch_ntt_internal<F: GoldilocksRepr>(inputs_ptr_in: *const F, outputs_ptr: *mut F, log_n: u32,
                                             num_ntts: u32, stride_between_input_arrays_in: u32,
                                             stride_between_output_arrays: u32, bitrev_inputs: bool, inverse: bool,
                                             log_extension_degree: u32, coset_index: u32, stream: &CudaStream)
                                             -> CudaResult<()> {
    check(unsafe {
        sys::ba_batch_ntt_internal(raw(inputs_ptr_in), raw_mut(outputs_ptr), log_n, num_ntts,
                                   stride_between_input_arrays_in, stride_between_output_arrays,
                                   bitrev_inputs as i32, inverse as i32, log_extension_degree, coset_index,
                                   stream.raw())
    }, "batch_ntt_internal")
}

#[allow(clippy::too_many_arguments)]
pub fn batch_ntt_out_of_place<F: GoldilocksRepr>(inputs_matrix: &DeviceSlice<F>, outputs_matrix: &mut DeviceSlice<F>,
                                                 log_n: u32, num_ntts: u32, inputs_offset: u32, outputs_offset: u32,
                                                 stride_between_input_arrays: u32,
                                                 stride_between_output_arrays: u32, bitrev_inputs: bool,
                                                 inverse: bool, log_extension_degree: u32, coset_index: u32,
                                                 stream: &CudaStream) -> CudaResult<()> {
    check(unsafe {
        sys::ba_batch_ntt_out_of_place(raw(inputs_matrix.as_ptr()), inputs_matrix.len(),
                                       raw_mut(outputs_matrix.as_mut_ptr()), outputs_matrix.len(), log_n, num_ntts,
                                       inputs_offset, outputs_offset, stride_between_input_arrays,
                                       stride_between_output_arrays, bitrev_inputs as i32, inverse as i32,
                                       log_extension_degree, coset_index, stream.raw())
    }, "batch_ntt_out_of_place")
}

#[allow(clippy::too_many_arguments)]
pub fn batch_ntt_in_place<F: GoldilocksRepr>(inputs_matrix: &mut DeviceSlice<F>, log_n: u32, num_ntts: u32,
                                             inputs_offset: u32, stride_between_input_arrays: u32,
                                             bitrev_inputs: bool, inverse: bool, log_extension_degree: u32,
                                             coset_index: u32, stream: &CudaStream) -> CudaResult<()> {
    check(unsafe {
        sys::ba_batch_ntt_in_place(raw_mut(inputs_matrix.as_mut_ptr()), inputs_matrix.len(), log_n, num_ntts,
                                   inputs_offset, stride_between_input_arrays, bitrev_inputs as i32, inverse as i32,
                                   log_extension_degree, coset_index, stream.raw())
    }, "batch_ntt_in_place")
}

/// Coset LDE of `num_cols` columns of 2^log_n evaluations (column k at k * in_stride) onto
/// 2^log_blowup cosets: coset c of column k at `out + k * out_col_stride + c * out_coset_stride`,
/// bit-reversed, cosets in shivini's order.  With `monomials` the evaluations are kept and the
/// monomials land there (bit-reversed, or natural with `natural_monomials`, as shivini's
/// `into_monomials` leaves them); without, they replace the evaluations.
#[allow(clippy::too_many_arguments)]
pub fn lde_batch<F: GoldilocksRepr>(ctx: &Context, evals: &mut DeviceSlice<F>, in_stride: u64,
                                    monomials: Option<&mut DeviceSlice<F>>, out: &mut DeviceSlice<F>,
                                    out_col_stride: u64, out_coset_stride: u64, log_n: u32, num_cols: u32,
                                    log_blowup: u32, stream: &CudaStream) -> CudaResult<()> {
    let n = 1u64 << log_n;
    let need_in = if num_cols == 0 { 0 } else { (num_cols as u64 - 1) * in_stride + n };
    let need_out = if num_cols == 0 { 0 } else {
        (num_cols as u64 - 1) * out_col_stride + ((1u64 << log_blowup) - 1) * out_coset_stride + n
    };
    assert!(evals.len() as u64 >= need_in && out.len() as u64 >= need_out, "lde_batch: buffers too small");
    let mono = match monomials {
        Some(m) => {
            assert!(m.len() as u64 >= n * num_cols as u64, "lde_batch: monomial buffer too small");
            raw_mut(m.as_mut_ptr())
        }
        None => core::ptr::null_mut(),
    };
    check(unsafe {
        sys::ba_lde_batch(ctx.raw(), raw_mut(evals.as_mut_ptr()), in_stride, mono, raw_mut(out.as_mut_ptr()),
                          out_col_stride, out_coset_stride, log_n, num_cols, log_blowup, stream.raw())
    }, "lde_batch")
}
