//! `CudaStream` (crates/cudart/src/stream.rs:17-132): `#[repr(transparent)]` over the handle,
//! the null handle being the default stream.  Streams are created non-blocking at the default
//! priority.
use crate::event::CudaEvent;
use crate::result::{CudaResult, CudaResultWrap};
use crate::sys;

#[repr(transparent)]
#[derive(Debug)]
pub struct CudaStream {
    handle: sys::ba_stream_t,
}

// the handle is an opaque runtime object usable from any thread, as in era_cudart
unsafe impl Send for CudaStream {}
unsafe impl Sync for CudaStream {}

#[derive(Debug, Clone, Copy, PartialEq, Eq, Default)]
pub struct CudaStreamCreateFlags(u32);
impl CudaStreamCreateFlags {
    pub const DEFAULT: Self = Self(0);
    pub const NON_BLOCKING: Self = Self(1);
}

#[derive(Debug, Clone, Copy, PartialEq, Eq, Default)]
pub struct CudaStreamWaitEventFlags(u32);
impl CudaStreamWaitEventFlags {
    pub const DEFAULT: Self = Self(0);
}

impl CudaStream {
    pub const DEFAULT: CudaStream = CudaStream { handle: core::ptr::null_mut() };

    pub fn create() -> CudaResult<Self> {
        let mut handle = core::ptr::null_mut();
        unsafe { sys::ba_stream_create(&mut handle) }.wrap_value(()).map(|_| Self { handle })
    }

    /// Every stream this crate creates is non-blocking; the flags are accepted for parity.
    pub fn create_with_flags(_flags: CudaStreamCreateFlags) -> CudaResult<Self> {
        Self::create()
    }

    pub fn destroy(self) -> CudaResult<()> {
        let handle = self.handle;
        core::mem::forget(self);
        if handle.is_null() {
            return Ok(());
        }
        unsafe { sys::ba_stream_destroy(handle) }.wrap()
    }

    pub fn synchronize(&self) -> CudaResult<()> {
        unsafe { sys::ba_stream_synchronize(self.handle) }.wrap()
    }

    pub fn wait_event(&self, event: &CudaEvent, _flags: CudaStreamWaitEventFlags) -> CudaResult<()> {
        unsafe { sys::ba_stream_wait_event(self.handle, event.raw()) }.wrap()
    }

    pub fn raw(&self) -> sys::ba_stream_t {
        self.handle
    }
}

impl Default for CudaStream {
    fn default() -> Self {
        Self::DEFAULT
    }
}

impl Drop for CudaStream {
    fn drop(&mut self) {
        if !self.handle.is_null() {
            unsafe { sys::ba_stream_destroy(self.handle) };
        }
    }
}
