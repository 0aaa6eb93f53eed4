//! Safe Rust mirror of boojum_cuda's NTT/LDE API over libboojum_amd.so.
//!
//! Module and function names, argument lists and types follow boojum_cuda (crates/boojum-cuda/
//! src/{context,ntt,ops_complex,barycentric,poseidon2}.rs) and era_cudart (through the
//! boojum-amd-cudart crate), so a caller that renames its dependencies
//! (`boojum_cuda = { package = "boojum-amd", features = ["boojum"] }`,
//! `era_cudart = { package = "boojum-amd-cudart" }`) keeps its call sites: the field-element
//! functions are generic over `GoldilocksRepr`, implemented for `u64` and, with the `boojum`
//! feature, for boojum's `GoldilocksField`; the stream argument is era_cudart's `&CudaStream`;
//! every call returns `CudaResult<()>`.  Status 1 from the library is where the reference
//! asserts: these wrappers keep that panic and return every other status as `Err`.
pub mod barycentric;
pub mod batch;
pub mod context;
pub mod ntt;
pub mod ops_complex;
pub mod poseidon2;
pub mod serve;

pub use boojum_amd_cudart as cudart;
pub use boojum_amd_sys as sys;

use cudart::result::{CudaError, CudaResult};

/// A Goldilocks field element stored as its `u64` representative: `#[repr(transparent)]` over
/// `u64` (boojum's `GoldilocksField` is), so device slices of it are the library's u64 arrays.
///
/// # Safety
/// Implementors are exactly 8 bytes, 8-aligned, with the value's u64 representative as their
/// only field.
pub unsafe trait GoldilocksRepr: Copy + 'static {}

unsafe impl GoldilocksRepr for u64 {}

#[cfg(feature = "boojum")]
unsafe impl GoldilocksRepr for boojum::field::goldilocks::GoldilocksField {}

/// Device pointer of a slice of field elements as the library's `u64` pointer.
pub(crate) fn raw<F: GoldilocksRepr>(p: *const F) -> *const u64 {
    assert!(core::mem::size_of::<F>() == 8 && core::mem::align_of::<F>() == 8);
    p as *const u64
}
pub(crate) fn raw_mut<F: GoldilocksRepr>(p: *mut F) -> *mut u64 {
    assert!(core::mem::size_of::<F>() == 8 && core::mem::align_of::<F>() == 8);
    p as *mut u64
}

/// Maps a status to a result; status 1 (invalid arguments) panics, as the reference's
/// `assert!`s do (e.g. crates/boojum-cuda/src/ntt.rs:192-200).
pub(crate) fn check(rc: i32, what: &str) -> CudaResult<()> {
    match rc {
        0 => Ok(()),
        1 => panic!("{what}: invalid arguments"),
        _ => Err(CudaError(rc)),
    }
}
