//! `CudaEvent` (crates/cudart/src/event.rs:16-121).
use crate::result::{CudaError, CudaResult, CudaResultWrap};
use crate::stream::CudaStream;
use crate::sys;

#[repr(transparent)]
#[derive(Debug)]
pub struct CudaEvent {
    handle: sys::ba_event_t,
}

unsafe impl Send for CudaEvent {}
unsafe impl Sync for CudaEvent {}

impl CudaEvent {
    pub fn create() -> CudaResult<Self> {
        let mut handle = core::ptr::null_mut();
        unsafe { sys::ba_event_create(&mut handle) }.wrap_value(()).map(|_| Self { handle })
    }

    pub fn record(&self, stream: &CudaStream) -> CudaResult<()> {
        unsafe { sys::ba_event_record(self.handle, stream.raw()) }.wrap()
    }

    pub fn synchronize(&self) -> CudaResult<()> {
        unsafe { sys::ba_event_synchronize(self.handle) }.wrap()
    }

    /// Ok(true) once complete, Ok(false) while pending.
    pub fn query(&self) -> CudaResult<bool> {
        match unsafe { sys::ba_event_query(self.handle) } {
            0 => Ok(true),
            600 => Ok(false),
            e => Err(CudaError(e)),
        }
    }

    pub fn raw(&self) -> sys::ba_event_t {
        self.handle
    }
}

/// `elapsed_time_in_ms` (crates/cudart/src/event.rs).
pub fn elapsed_time_in_ms(start: &CudaEvent, end: &CudaEvent) -> CudaResult<f32> {
    let mut ms = 0f32;
    unsafe { sys::ba_event_elapsed_ms(&mut ms, start.raw(), end.raw()) }.wrap_value(()).map(|_| ms)
}

impl Drop for CudaEvent {
    fn drop(&mut self) {
        unsafe { sys::ba_event_destroy(self.handle) };
    }
}
