//! `boojum_cuda::ops_complex` (crates/boojum-cuda/src/ops_complex.rs:31-331): bit reversal of
//! column-major matrices, vectors of powers and omega shifts.  `result.len()` is the count, as
//! in the reference; extension-field results hold two planes of `count` elements.
use crate::cudart::result::CudaResult;
use crate::cudart::slice::DeviceSlice;
use crate::cudart::stream::CudaStream;
use crate::{check, raw, raw_mut, sys, GoldilocksRepr};

pub fn bit_reverse<F: GoldilocksRepr>(src: &DeviceSlice<F>, src_stride: usize, dst: &mut DeviceSlice<F>, dst_stride: usize,
                   log_rows: u32, cols: u32, stream: &CudaStream) -> CudaResult<()> {
    check(unsafe {
        sys::ba_bit_reverse(raw(src.as_ptr()), src_stride as u64, raw_mut(dst.as_mut_ptr()), dst_stride as u64, log_rows, cols,
                            stream.raw())
    }, "bit_reverse")
}

pub fn bit_reverse_in_place<F: GoldilocksRepr>(values: &mut DeviceSlice<F>, stride: usize, log_rows: u32, cols: u32,
                            stream: &CudaStream) -> CudaResult<()> {
    check(unsafe { sys::ba_bit_reverse_in_place(raw_mut(values.as_mut_ptr()), stride as u64, log_rows, cols, stream.raw()) },
          "bit_reverse_in_place")
}

pub fn get_powers_of_w<F: GoldilocksRepr>(log_degree: u32, offset: u32, inverse: bool, bit_reverse: bool,
                       result: &mut DeviceSlice<F>,
                       stream: &CudaStream) -> CudaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_of_w(log_degree, offset, inverse as i32, bit_reverse as i32, raw_mut(result.as_mut_ptr()),
                                result.len() as u32, stream.raw())
    }, "get_powers_of_w")
}

pub fn get_powers_of_g<F: GoldilocksRepr>(log_degree: u32, offset: u32, inverse: bool, bit_reverse: bool,
                       result: &mut DeviceSlice<F>,
                       stream: &CudaStream) -> CudaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_of_g(log_degree, offset, inverse as i32, bit_reverse as i32, raw_mut(result.as_mut_ptr()),
                                result.len() as u32, stream.raw())
    }, "get_powers_of_g")
}

/// Base-field `get_powers_by_val`.
pub fn get_powers_by_val<F: GoldilocksRepr>(base: u64, offset: u32, bit_reverse: bool, result: &mut DeviceSlice<F>,
                         stream: &CudaStream) -> CudaResult<()> {
    assert!(result.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_by_val_bf(base, offset, bit_reverse as i32, raw_mut(result.as_mut_ptr()), result.len() as u32,
                                     stream.raw())
    }, "get_powers_by_val")
}

/// Extension-field `get_powers_by_val`: `result` holds 2 * count elements (planes c0, c1).
pub fn get_powers_by_val_ext(base: [u64; 2], offset: u32, bit_reverse: bool, result: &mut DeviceSlice<u64>,
                             stream: &CudaStream) -> CudaResult<()> {
    assert!(result.len() % 2 == 0 && result.len() / 2 <= u32::MAX as usize);
    check(unsafe {
        sys::ba_get_powers_by_val_ef(base[0], base[1], offset, bit_reverse as i32, result.as_mut_ptr(),
                                     (result.len() / 2) as u32, stream.raw())
    }, "get_powers_by_val")
}

pub fn omega_shift<F: GoldilocksRepr>(src: &DeviceSlice<F>, log_degree: u32, offset: u32, inverse: bool, shift: u32,
                   dst: &mut DeviceSlice<F>, stream: &CudaStream) -> CudaResult<()> {
    assert_eq!(src.len(), dst.len());
    assert!(dst.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_omega_shift(raw(src.as_ptr()), log_degree, offset, inverse as i32, shift, raw_mut(dst.as_mut_ptr()),
                            dst.len() as u32, stream.raw())
    }, "omega_shift")
}

pub fn omega_shift_in_place<F: GoldilocksRepr>(values: &mut DeviceSlice<F>, log_degree: u32, offset: u32, inverse: bool, shift: u32,
                            stream: &CudaStream) -> CudaResult<()> {
    assert!(values.len() <= u32::MAX as usize);
    check(unsafe {
        sys::ba_omega_shift_in_place(raw_mut(values.as_mut_ptr()), log_degree, offset, inverse as i32, shift,
                                     values.len() as u32, stream.raw())
    }, "omega_shift_in_place")
}
