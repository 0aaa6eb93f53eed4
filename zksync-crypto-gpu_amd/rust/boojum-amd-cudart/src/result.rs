//! `CudaResult` / `CudaError` (crates/cudart/src/result.rs:5, crates/cudart-sys/src/bindings.rs:30).
//! The error is the library's int status: 0 or a HIP error code (hipErrorInvalidValue = 1,
//! hipErrorOutOfMemory = 2, hipErrorNotInitialized = 3, ...).
use crate::sys;

#[repr(transparent)]
#[derive(Clone, Copy, Debug, PartialEq, Eq, Hash)]
pub struct CudaError(pub i32);

#[allow(non_upper_case_globals)]
impl CudaError {
    pub const Success: CudaError = CudaError(sys::BA_SUCCESS);
    pub const ErrorInvalidValue: CudaError = CudaError(sys::BA_ERROR_INVALID_VALUE);
    pub const ErrorMemoryAllocation: CudaError = CudaError(sys::BA_ERROR_OUT_OF_MEMORY);
    pub const ErrorInitializationError: CudaError = CudaError(sys::BA_ERROR_NOT_INITIALIZED);
    pub const ErrorNotReady: CudaError = CudaError(600);
}

impl core::fmt::Display for CudaError {
    fn fmt(&self, f: &mut core::fmt::Formatter<'_>) -> core::fmt::Result {
        let name = unsafe { std::ffi::CStr::from_ptr(sys::ba_error_name(self.0)) };
        write!(f, "status {} ({})", self.0, name.to_string_lossy())
    }
}

impl std::error::Error for CudaError {}

pub type CudaResult<T> = Result<T, CudaError>;

/// `CudaResultWrap` (crates/cudart/src/result.rs:7-29) for the library's int statuses.
pub trait CudaResultWrap {
    fn wrap(self) -> CudaResult<()>;
    fn wrap_value<T>(self, value: T) -> CudaResult<T>;
}

impl CudaResultWrap for i32 {
    fn wrap(self) -> CudaResult<()> {
        self.wrap_value(())
    }
    fn wrap_value<T>(self, value: T) -> CudaResult<T> {
        if self == sys::BA_SUCCESS {
            Ok(value)
        } else {
            Err(CudaError(self))
        }
    }
}
