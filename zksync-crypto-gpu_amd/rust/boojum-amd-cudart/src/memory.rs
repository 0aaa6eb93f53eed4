//! Device allocations and copies (crates/cudart/src/memory.rs).
use core::ffi::c_void;
use core::marker::PhantomData;
use core::ops::{Deref, DerefMut};

use crate::result::{CudaResult, CudaResultWrap};
use crate::slice::{CudaSlice, CudaSliceMut, DeviceSlice};
use crate::stream::CudaStream;
use crate::sys;

/// hipMemcpyDefault: the runtime infers the direction from the pointers (unified addressing).
const COPY_DEFAULT: i32 = 4;

/// Owned device memory (`DeviceAllocation<T>`), dereferencing to a `DeviceSlice<T>`.
pub struct DeviceAllocation<T> {
    ptr: *mut T,
    len: usize,
    _t: PhantomData<T>,
}

unsafe impl<T: Send> Send for DeviceAllocation<T> {}
unsafe impl<T: Sync> Sync for DeviceAllocation<T> {}

impl<T> DeviceAllocation<T> {
    pub fn alloc(length: usize) -> CudaResult<Self> {
        let mut p: *mut c_void = core::ptr::null_mut();
        unsafe { sys::ba_malloc(&mut p, length * core::mem::size_of::<T>()) }
            .wrap_value(())
            .map(|_| Self { ptr: p as *mut T, len: length, _t: PhantomData })
    }
    pub fn free(self) -> CudaResult<()> {
        let p = self.ptr;
        core::mem::forget(self);
        unsafe { sys::ba_free(p as *mut c_void) }.wrap()
    }
}

impl<T> Deref for DeviceAllocation<T> {
    type Target = DeviceSlice<T>;
    fn deref(&self) -> &DeviceSlice<T> {
        unsafe { DeviceSlice::from_raw_parts(self.ptr, self.len) }
    }
}

impl<T> DerefMut for DeviceAllocation<T> {
    fn deref_mut(&mut self) -> &mut DeviceSlice<T> {
        unsafe { DeviceSlice::from_raw_parts_mut(self.ptr, self.len) }
    }
}

impl<T> Drop for DeviceAllocation<T> {
    fn drop(&mut self) {
        unsafe { sys::ba_free(self.ptr as *mut c_void) };
    }
}

/// `memory_copy_async` between any host / device slices of equal length.
pub fn memory_copy_async<T, Dst, Src>(dst: &mut Dst, src: &Src, stream: &CudaStream) -> CudaResult<()>
where
    Dst: CudaSliceMut<T> + ?Sized,
    Src: CudaSlice<T> + ?Sized,
{
    assert_eq!(dst.c_len(), src.c_len());
    let bytes = src.c_len() * core::mem::size_of::<T>();
    unsafe {
        sys::ba_memcpy_async(dst.as_mut_c_ptr() as *mut c_void, src.as_c_ptr() as *const c_void, bytes, COPY_DEFAULT,
                             stream.raw())
    }
    .wrap()
}

/// `memory_copy` (synchronous).
pub fn memory_copy<T, Dst, Src>(dst: &mut Dst, src: &Src) -> CudaResult<()>
where
    Dst: CudaSliceMut<T> + ?Sized,
    Src: CudaSlice<T> + ?Sized,
{
    assert_eq!(dst.c_len(), src.c_len());
    let bytes = src.c_len() * core::mem::size_of::<T>();
    unsafe { sys::ba_memcpy(dst.as_mut_c_ptr() as *mut c_void, src.as_c_ptr() as *const c_void, bytes, COPY_DEFAULT) }
        .wrap()
}
