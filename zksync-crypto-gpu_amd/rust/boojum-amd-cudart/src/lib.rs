//! The era_cudart subset the Goldilocks NTT path rides on, over libboojum_amd.so's HIP
//! pass-throughs.  Module paths and type names follow era_cudart (crates/cudart/src/lib.rs) so
//! that a crate depending on `era_cudart = { package = "boojum-amd-cudart", ... }` keeps its
//! `use era_cudart::{result, stream, event, memory, slice}` lines.
pub mod device;
pub mod event;
pub mod memory;
pub mod result;
pub mod slice;
pub mod stream;

pub use boojum_amd_sys as sys;
