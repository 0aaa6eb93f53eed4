//! Safe Rust mirror of the reference's NTT/LDE API over libboojum_amd.so.
//!
//! Module and function names follow the reference crates so a caller switches by changing
//! its `use` lines: `runtime` stands for era_cudart (crates/cudart/src/{stream,event,memory,
//! slice}), `context`, `ntt`, `ops_complex`, `barycentric` and `poseidon2` for boojum_cuda's
//! modules of the same names.  Field elements are `u64` (boojum's `GoldilocksField` is `#[repr(transparent)]`
//! over `u64`, so a caller's `DeviceSlice<GoldilocksField>` is viewed as `DeviceSlice<u64>`).
//! Status 1 from the library is where the reference asserts: these wrappers keep that panic,
//! and return every other error as `Err`.
pub mod barycentric;
pub mod batch;
pub mod context;
pub mod ntt;
pub mod ops_complex;
pub mod poseidon2;
pub mod runtime;
pub mod serve;

pub use boojum_amd_sys as sys;

/// A non-zero status of the library (a HIP error code or one of `sys::BA_ERROR_*`).
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct BaError(pub i32);

impl core::fmt::Display for BaError {
    fn fmt(&self, f: &mut core::fmt::Formatter<'_>) -> core::fmt::Result {
        let name = unsafe { std::ffi::CStr::from_ptr(sys::ba_error_name(self.0)) };
        write!(f, "boojum-amd status {} ({})", self.0, name.to_string_lossy())
    }
}

impl std::error::Error for BaError {}

pub type BaResult<T> = Result<T, BaError>;

/// Maps a status to a result; status 1 (invalid arguments) panics, as the reference's
/// `assert!`s do (e.g. crates/boojum-cuda/src/ntt.rs:192-200).
pub(crate) fn check(rc: i32, what: &str) -> BaResult<()> {
    match rc {
        0 => Ok(()),
        1 => panic!("{what}: invalid arguments"),
        _ => Err(BaError(rc)),
    }
}
