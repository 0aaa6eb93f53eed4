//! `boojum_cuda::ntt` (crates/boojum-cuda/src/ntt.rs:179-369), same argument lists; plus the
//! LDE shivini composes from it (crates/shivini/src/data_structures/storage.rs:154-260).
use crate::context::Context;
use crate::runtime::{CudaStream, DeviceSlice};
use crate::{check, sys, BaResult};

/// # Safety
/// `inputs` / `outputs` are device pointers covering `num_ntts` strided arrays of 2^log_n.
#[allow(clippy::too_many_arguments)]
pub unsafe fn batch_ntt_internal(inputs: *const u64, outputs: *mut u64, log_n: u32, num_ntts: u32,
                                 stride_between_input_arrays: u32, stride_between_output_arrays: u32,
                                 bitrev_inputs: bool, inverse: bool, log_extension_degree: u32, coset_index: u32,
                                 stream: &CudaStream) -> BaResult<()> {
    check(sys::ba_batch_ntt_internal(inputs, outputs, log_n, num_ntts, stride_between_input_arrays,
                                     stride_between_output_arrays, bitrev_inputs as i32, inverse as i32,
                                     log_extension_degree, coset_index, stream.raw()), "batch_ntt_internal")
}

#[allow(clippy::too_many_arguments)]
pub fn batch_ntt_out_of_place(inputs_matrix: &DeviceSlice<u64>, outputs_matrix: &mut DeviceSlice<u64>, log_n: u32,
                              num_ntts: u32, inputs_offset: u32, outputs_offset: u32,
                              stride_between_input_arrays: u32, stride_between_output_arrays: u32,
                              bitrev_inputs: bool, inverse: bool, log_extension_degree: u32, coset_index: u32,
                              stream: &CudaStream) -> BaResult<()> {
    check(unsafe {
        sys::ba_batch_ntt_out_of_place(inputs_matrix.as_ptr(), inputs_matrix.len(), outputs_matrix.as_mut_ptr(),
                                       outputs_matrix.len(), log_n, num_ntts, inputs_offset, outputs_offset,
                                       stride_between_input_arrays, stride_between_output_arrays,
                                       bitrev_inputs as i32, inverse as i32, log_extension_degree, coset_index,
                                       stream.raw())
    }, "batch_ntt_out_of_place")
}

#[allow(clippy::too_many_arguments)]
pub fn batch_ntt_in_place(inputs_matrix: &mut DeviceSlice<u64>, log_n: u32, num_ntts: u32, inputs_offset: u32,
                          stride_between_input_arrays: u32, bitrev_inputs: bool, inverse: bool,
                          log_extension_degree: u32, coset_index: u32, stream: &CudaStream) -> BaResult<()> {
    check(unsafe {
        sys::ba_batch_ntt_in_place(inputs_matrix.as_mut_ptr(), inputs_matrix.len(), log_n, num_ntts, inputs_offset,
                                   stride_between_input_arrays, bitrev_inputs as i32, inverse as i32,
                                   log_extension_degree, coset_index, stream.raw())
    }, "batch_ntt_in_place")
}

/// Coset LDE of `num_cols` columns of 2^log_n evaluations (column k at k * in_stride) onto
/// 2^log_blowup cosets: coset c of column k at `out + k * out_col_stride + c * out_coset_stride`,
/// bit-reversed, cosets in shivini's order.  With `monomials` the evaluations are kept and the
/// bit-reversed monomials land there; without, they replace the evaluations.
#[allow(clippy::too_many_arguments)]
pub fn lde_batch(ctx: &Context, evals: &mut DeviceSlice<u64>, in_stride: u64, monomials: Option<&mut DeviceSlice<u64>>,
                 out: &mut DeviceSlice<u64>, out_col_stride: u64, out_coset_stride: u64, log_n: u32, num_cols: u32,
                 log_blowup: u32, stream: &CudaStream) -> BaResult<()> {
    let n = 1u64 << log_n;
    let need_in = if num_cols == 0 { 0 } else { (num_cols as u64 - 1) * in_stride + n };
    let need_out = if num_cols == 0 { 0 } else {
        (num_cols as u64 - 1) * out_col_stride + ((1u64 << log_blowup) - 1) * out_coset_stride + n
    };
    assert!(evals.len() as u64 >= need_in && out.len() as u64 >= need_out, "lde_batch: buffers too small");
    let mono = match monomials {
        Some(m) => {
            assert!(m.len() as u64 >= n * num_cols as u64, "lde_batch: monomial buffer too small");
            m.as_mut_ptr()
        }
        None => core::ptr::null_mut(),
    };
    check(unsafe {
        sys::ba_lde_batch(ctx.raw(), evals.as_mut_ptr(), in_stride, mono, out.as_mut_ptr(), out_col_stride,
                          out_coset_stride, log_n, num_cols, log_blowup, stream.raw())
    }, "lde_batch")
}
