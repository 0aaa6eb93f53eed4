//! `boojum_cuda::ops_complex` (crates/boojum-cuda/src/ops_complex.rs:31-331): bit reversal of
//! column-major matrices, vectors of powers and omega shifts.  `result.len()` is the count, as
//! in the reference; extension-field results hold two planes of `count` elements.
use crate::runtime::{CudaStream, DeviceSlice};
use crate::{check, sys, BaResult};

pub fn bit_reverse(src: &DeviceSlice<u64>, src_stride: usize, dst: &mut DeviceSlice<u64>, dst_stride: usize,
                   log_rows: u32, cols: u32, stream: &CudaStream) -> BaResult<()> {
    check(unsafe {
        sys::ba_bit_reverse(src.as_ptr(), src_stride as u64, dst.as_mut_ptr(), dst_stride as u64, log_rows, cols,
                            stream.raw())
    }, "bit_reverse")
}

pub fn bit_reverse_in_place(values: &mut DeviceSlice<u64>, stride: usize, log_rows: u32, cols: u32,
                            stream: &CudaStream) -> BaResult<()> {
    check(unsafe { sys::ba_bit_reverse_in_place(values.as_mut_ptr(), stride as u64, log_rows, cols, stream.raw()) },
          "bit_reverse_in_place")
}

pub fn get_powers_of_w(log_degree: u32, offset: u32, inverse: bool, bit_reverse: bool, result: &mut DeviceSlice<u64>,
                       stream: &CudaStream) -> BaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_of_w(log_degree, offset, inverse as i32, bit_reverse as i32, result.as_mut_ptr(),
                                result.len() as u32, stream.raw())
    }, "get_powers_of_w")
}

pub fn get_powers_of_g(log_degree: u32, offset: u32, inverse: bool, bit_reverse: bool, result: &mut DeviceSlice<u64>,
                       stream: &CudaStream) -> BaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_of_g(log_degree, offset, inverse as i32, bit_reverse as i32, result.as_mut_ptr(),
                                result.len() as u32, stream.raw())
    }, "get_powers_of_g")
}

/// Base-field `get_powers_by_val`.
pub fn get_powers_by_val(base: u64, offset: u32, bit_reverse: bool, result: &mut DeviceSlice<u64>,
                         stream: &CudaStream) -> BaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_by_val_bf(base, offset, bit_reverse as i32, result.as_mut_ptr(), result.len() as u32,
                                     stream.raw())
    }, "get_powers_by_val")
}

/// Extension-field `get_powers_by_val`: `result` holds 2 * count elements (planes c0, c1).
pub fn get_powers_by_val_ext(base: [u64; 2], offset: u32, bit_reverse: bool, result: &mut DeviceSlice<u64>,
                             stream: &CudaStream) -> BaResult<()> {
    assert!(result.len() % 2 == 0 && result.len() / 2 <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_by_val_ef(base[0], base[1], offset, bit_reverse as i32, result.as_mut_ptr(),
                                     (result.len() / 2) as u32, stream.raw())
    }, "get_powers_by_val")
}

pub fn omega_shift(src: &DeviceSlice<u64>, log_degree: u32, offset: u32, inverse: bool, shift: u32,
                   dst: &mut DeviceSlice<u64>, stream: &CudaStream) -> BaResult<()> {
    assert_eq!(src.len(), dst.len());
    assert!(dst.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_omega_shift(src.as_ptr(), log_degree, offset, inverse as i32, shift, dst.as_mut_ptr(),
                            dst.len() as u32, stream.raw())
    }, "omega_shift")
}

pub fn omega_shift_in_place(values: &mut DeviceSlice<u64>, log_degree: u32, offset: u32, inverse: bool, shift: u32,
                            stream: &CudaStream) -> BaResult<()> {
    assert!(values.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_omega_shift_in_place(values.as_mut_ptr(), log_degree, offset, inverse as i32, shift,
                                     values.len() as u32, stream.raw())
    }, "omega_shift_in_place")
}
