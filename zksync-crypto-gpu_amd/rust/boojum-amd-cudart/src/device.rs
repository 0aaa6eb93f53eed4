//! Device selection (crates/cudart/src/device.rs).
use crate::result::{CudaResult, CudaResultWrap};
use crate::sys;

pub fn get_device_count() -> CudaResult<i32> {
    let mut n = 0;
    unsafe { sys::ba_get_device_count(&mut n) }.wrap_value(()).map(|_| n)
}

pub fn get_device() -> CudaResult<i32> {
    let mut d = 0;
    unsafe { sys::ba_get_device(&mut d) }.wrap_value(()).map(|_| d)
}

pub fn set_device(device_id: i32) -> CudaResult<()> {
    unsafe { sys::ba_set_device(device_id) }.wrap()
}

pub fn device_synchronize() -> CudaResult<()> {
    unsafe { sys::ba_device_synchronize() }.wrap()
}
