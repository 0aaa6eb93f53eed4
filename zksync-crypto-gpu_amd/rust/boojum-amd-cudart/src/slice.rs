//! `DeviceSlice<T>` (crates/cudart/src/slice/device_slice.rs:11-80): `#[repr(transparent)]` over
//! `[T]`, so it carries pointer and length; its elements live on the device and are never read
//! or written on the host.  `CudaSlice` / `CudaSliceMut` (crates/cudart/src/slice/mod.rs) let
//! copies take host and device slices alike.
#[repr(transparent)]
pub struct DeviceSlice<T>([T]);

impl<T> DeviceSlice<T> {
    /// # Safety
    /// `data_address` must point to `len` elements of device memory that outlive `'a`.
    pub unsafe fn from_raw_parts<'a>(data_address: *const T, len: usize) -> &'a Self {
        &*(core::ptr::slice_from_raw_parts(data_address, len) as *const Self)
    }
    /// # Safety
    /// As `from_raw_parts`, and no other reference to the memory is live.
    pub unsafe fn from_raw_parts_mut<'a>(data_address: *mut T, len: usize) -> &'a mut Self {
        &mut *(core::ptr::slice_from_raw_parts_mut(data_address, len) as *mut Self)
    }
    /// # Safety
    /// The slice must point to device memory.
    pub unsafe fn from_slice(slice: &[T]) -> &Self {
        Self::from_raw_parts(slice.as_ptr(), slice.len())
    }
    /// # Safety
    /// The slice must point to device memory.
    pub unsafe fn from_mut_slice(slice: &mut [T]) -> &mut Self {
        Self::from_raw_parts_mut(slice.as_mut_ptr(), slice.len())
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
    /// # Safety
    /// The returned slice aliases device memory: it must not be read or written on the host.
    pub unsafe fn as_slice(&self) -> &[T] {
        &self.0
    }
    /// # Safety
    /// As `as_slice`.
    pub unsafe fn as_mut_slice(&mut self) -> &mut [T] {
        &mut self.0
    }
    pub fn split_at(&self, mid: usize) -> (&Self, &Self) {
        let (a, b) = self.0.split_at(mid);
        unsafe { (Self::from_slice(a), Self::from_slice(b)) }
    }
    pub fn split_at_mut(&mut self, mid: usize) -> (&mut Self, &mut Self) {
        let (a, b) = self.0.split_at_mut(mid);
        unsafe { (Self::from_mut_slice(a), Self::from_mut_slice(b)) }
    }
}

/// Memory a copy can read: host slices and device slices.
pub trait CudaSlice<T> {
    fn as_c_ptr(&self) -> *const T;
    fn c_len(&self) -> usize;
}
/// Memory a copy can write.
pub trait CudaSliceMut<T>: CudaSlice<T> {
    fn as_mut_c_ptr(&mut self) -> *mut T;
}

impl<T> CudaSlice<T> for [T] {
    fn as_c_ptr(&self) -> *const T {
        self.as_ptr()
    }
    fn c_len(&self) -> usize {
        self.len()
    }
}
impl<T> CudaSliceMut<T> for [T] {
    fn as_mut_c_ptr(&mut self) -> *mut T {
        self.as_mut_ptr()
    }
}
impl<T> CudaSlice<T> for DeviceSlice<T> {
    fn as_c_ptr(&self) -> *const T {
        self.as_ptr()
    }
    fn c_len(&self) -> usize {
        self.len()
    }
}
impl<T> CudaSliceMut<T> for DeviceSlice<T> {
    fn as_mut_c_ptr(&mut self) -> *mut T {
        self.as_mut_ptr()
    }
}
