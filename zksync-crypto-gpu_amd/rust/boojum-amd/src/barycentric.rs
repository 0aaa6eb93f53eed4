//! `boojum_cuda::barycentric` (crates/boojum-cuda/src/barycentric.rs:83-300).
use crate::cudart::result::CudaResult;
use crate::cudart::slice::DeviceSlice;
use crate::cudart::stream::CudaStream;
use crate::{check, sys};

/// Where the point X lies, and the evaluations' field, as the reference's Precompute/Eval impls.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Eval {
    BaseAtBase = 0,
    BaseAtExt = 1,
    ExtAtExt = 2,
}

/// `precompute_lagrange_coeffs::<PrecomputeAtBase | PrecomputeAtExt>`: `x` and
/// `common_factor_storage` hold 1 (base) or 2 (extension) elements; `lagrange_coeffs` holds
/// count (base) or 2 * count (extension planes) elements, count a power of two.
pub fn precompute_lagrange_coeffs(ext: bool, x: &DeviceSlice<u64>, common_factor_storage: &mut DeviceSlice<u64>,
                                  coset: u64, lagrange_coeffs: &mut DeviceSlice<u64>, bit_reverse: bool,
                                  stream: &CudaStream) -> CudaResult<()> {
    let width = if ext { 2 } else { 1 };
    assert!(x.len() >= width && common_factor_storage.len() >= width);
    let count = lagrange_coeffs.len() / width;
    assert!(count.is_power_of_two() && count <= u32::MAX as usize);
    check(unsafe {
        sys::ba_barycentric_precompute_lagrange_coeffs(ext as i32, x.as_ptr(), common_factor_storage.as_mut_ptr(),
                                                       coset, lagrange_coeffs.as_mut_ptr(), count as u32,
                                                       bit_reverse as i32, stream.raw())
    }, "precompute_lagrange_coeffs")
}

/// Elements of temp storage `batch_eval` needs (the reference's get_batch_eval_temp_storage_sizes).
pub fn get_batch_eval_temp_storage_elems(kind: Eval, count: u32, num_polys: u32) -> CudaResult<usize> {
    let mut n = 0u64;
    check(unsafe { sys::ba_barycentric_batch_eval_temp_elems(kind as i32, count, num_polys, &mut n) },
          "batch_eval_temp")?;
    Ok(n as usize)
}

/// `batch_eval`: evals holds num_polys base elements (BaseAtBase) or (c0, c1) pairs.
#[allow(clippy::too_many_arguments)]
pub fn batch_eval(kind: Eval, batch_ys: &DeviceSlice<u64>, count: u32, num_polys: u32,
                  lagrange_coeffs: &DeviceSlice<u64>, temp: &mut DeviceSlice<u64>, evals: &mut DeviceSlice<u64>,
                  stream: &CudaStream) -> CudaResult<()> {
    let width = if kind == Eval::BaseAtBase { 1 } else { 2 };
    assert_eq!(evals.len(), width * num_polys as usize);
    check(unsafe {
        sys::ba_barycentric_batch_eval(kind as i32, batch_ys.as_ptr(), count, num_polys, lagrange_coeffs.as_ptr(),
                                       temp.as_mut_ptr(), temp.len() as u64, evals.as_mut_ptr(), stream.raw())
    }, "batch_eval")
}
