//! The era_cudart subset the NTT path uses: streams, events, device memory, device slices.
use crate::{check, sys, BaResult};
use core::ffi::c_void;
use core::marker::PhantomData;

/// A stream (`CudaStream`, crates/cudart/src/stream.rs:17-132); the default is the null stream.
pub struct CudaStream {
    pub(crate) handle: sys::ba_stream_t,
    owned: bool,
}

impl CudaStream {
    pub fn create() -> BaResult<Self> {
        let mut h = core::ptr::null_mut();
        check(unsafe { sys::ba_stream_create(&mut h) }, "stream_create")?;
        Ok(Self { handle: h, owned: true })
    }
    pub fn synchronize(&self) -> BaResult<()> {
        check(unsafe { sys::ba_stream_synchronize(self.handle) }, "stream_synchronize")
    }
    pub fn wait_event(&self, event: &CudaEvent) -> BaResult<()> {
        check(unsafe { sys::ba_stream_wait_event(self.handle, event.handle) }, "stream_wait_event")
    }
    pub fn raw(&self) -> sys::ba_stream_t {
        self.handle
    }
}

impl Default for CudaStream {
    fn default() -> Self {
        Self { handle: core::ptr::null_mut(), owned: false }
    }
}

impl Drop for CudaStream {
    fn drop(&mut self) {
        if self.owned {
            unsafe { sys::ba_stream_destroy(self.handle) };
        }
    }
}

/// An event (`CudaEvent`, crates/cudart/src/event.rs:16-121).
pub struct CudaEvent {
    pub(crate) handle: sys::ba_event_t,
}

impl CudaEvent {
    pub fn create() -> BaResult<Self> {
        let mut h = core::ptr::null_mut();
        check(unsafe { sys::ba_event_create(&mut h) }, "event_create")?;
        Ok(Self { handle: h })
    }
    pub fn record(&self, stream: &CudaStream) -> BaResult<()> {
        check(unsafe { sys::ba_event_record(self.handle, stream.handle) }, "event_record")
    }
    pub fn synchronize(&self) -> BaResult<()> {
        check(unsafe { sys::ba_event_synchronize(self.handle) }, "event_synchronize")
    }
    pub fn elapsed_ms(&self, end: &CudaEvent) -> BaResult<f32> {
        let mut ms = 0f32;
        check(unsafe { sys::ba_event_elapsed_ms(&mut ms, self.handle, end.handle) }, "event_elapsed")?;
        Ok(ms)
    }
}

impl Drop for CudaEvent {
    fn drop(&mut self) {
        unsafe { sys::ba_event_destroy(self.handle) };
    }
}

/// A slice of device memory (`DeviceSlice<T>`, crates/cudart/src/slice/device_slice.rs:11-80):
/// `#[repr(transparent)]` over `[T]` so it carries pointer and length; its elements live on the
/// device and are never read or written on the host.
#[repr(transparent)]
pub struct DeviceSlice<T>([T]);

impl<T> DeviceSlice<T> {
    /// # Safety
    /// `ptr` must point to `len` elements of device memory that outlive `'a`.
    pub unsafe fn from_raw_parts<'a>(ptr: *const T, len: usize) -> &'a Self {
        &*(core::ptr::slice_from_raw_parts(ptr, len) as *const Self)
    }
    /// # Safety
    /// As `from_raw_parts`, and no other reference to the memory is live.
    pub unsafe fn from_raw_parts_mut<'a>(ptr: *mut T, len: usize) -> &'a mut Self {
        &mut *(core::ptr::slice_from_raw_parts_mut(ptr, len) as *mut Self)
    }
    pub fn len(&self) -> usize {
        self.0.len()
    }
    pub fn is_empty(&self) -> bool {
        self.0.is_empty()
    }
    pub fn as_ptr(&self) -> *const T {
        self.0.as_ptr()
    }
    pub fn as_mut_ptr(&mut self) -> *mut T {
        self.0.as_mut_ptr()
    }
}

/// Owned device memory (`DeviceAllocation<T>`, crates/cudart/src/memory.rs:17).
pub struct DeviceAllocation<T> {
    ptr: *mut T,
    len: usize,
    _t: PhantomData<T>,
}

impl<T> DeviceAllocation<T> {
    pub fn alloc(len: usize) -> BaResult<Self> {
        let mut p: *mut c_void = core::ptr::null_mut();
        check(unsafe { sys::ba_malloc(&mut p, len * core::mem::size_of::<T>()) }, "malloc")?;
        Ok(Self { ptr: p as *mut T, len, _t: PhantomData })
    }
    pub fn as_slice(&self) -> &DeviceSlice<T> {
        unsafe { DeviceSlice::from_raw_parts(self.ptr, self.len) }
    }
    pub fn as_mut_slice(&mut self) -> &mut DeviceSlice<T> {
        unsafe { DeviceSlice::from_raw_parts_mut(self.ptr, self.len) }
    }
}

impl<T> Drop for DeviceAllocation<T> {
    fn drop(&mut self) {
        unsafe { sys::ba_free(self.ptr as *mut c_void) };
    }
}

const H2D: i32 = 1;
const D2H: i32 = 2;

/// `memory_copy_async` from host to device (crates/cudart/src/memory.rs).
pub fn memory_copy_to_device_async<T: Copy>(dst: &mut DeviceSlice<T>, src: &[T], stream: &CudaStream) -> BaResult<()> {
    assert_eq!(dst.len(), src.len());
    let bytes = src.len() * core::mem::size_of::<T>();
    check(unsafe { sys::ba_memcpy_async(dst.as_mut_ptr() as *mut c_void, src.as_ptr() as *const c_void, bytes, H2D,
                                        stream.handle) }, "memcpy")
}

/// `memory_copy_async` from device to host.
pub fn memory_copy_to_host_async<T: Copy>(dst: &mut [T], src: &DeviceSlice<T>, stream: &CudaStream) -> BaResult<()> {
    assert_eq!(dst.len(), src.len());
    let bytes = src.len() * core::mem::size_of::<T>();
    check(unsafe { sys::ba_memcpy_async(dst.as_mut_ptr() as *mut c_void, src.as_ptr() as *const c_void, bytes, D2H,
                                        stream.handle) }, "memcpy")
}
