//! `boojum_cuda::poseidon2::GpuTreeHasher` for the Goldilocks Poseidon2 hasher
//! (crates/boojum-cuda/src/poseidon2.rs:165-316).  Digests are 4 planes of `u64`: word k of digest
//! g at `results[g + k * count]`.  The context needs boojum's round constants first
//! (`set_round_constants`, e.g. with `poseidon_goldilocks_params::ALL_ROUND_CONSTANTS_AS_FIELD_ELEMENTS`
//! viewed as `u64`).
use crate::context::Context;
use crate::cudart::result::CudaResult;
use crate::cudart::slice::DeviceSlice;
use crate::cudart::stream::CudaStream;
use crate::{check, sys};

pub const RATE: usize = 8;
pub const CAPACITY: usize = 4;
pub const ROUNDS: usize = 30;
pub const WIDTH: usize = 12;

/// Installs the 30 x 12 round constants (row-major by round) into `ctx`.
pub fn set_round_constants(ctx: &Context, round_constants: &[u64; ROUNDS * WIDTH]) -> CudaResult<()> {
    check(unsafe { sys::ba_poseidon2_gl_set_round_constants(ctx.raw(), round_constants.as_ptr()) },
          "poseidon2 set_round_constants")
}

pub fn compute_leaf_hashes(values: &DeviceSlice<u64>, results: &mut DeviceSlice<u64>, log_rows_per_hash: u32,
                           load_intermediate: bool, store_intermediate: bool, stream: &CudaStream) -> CudaResult<()> {
    check(unsafe {
        sys::ba_poseidon2_gl_compute_leaf_hashes(values.as_ptr(), values.len(), results.as_mut_ptr(), results.len(),
                                                 log_rows_per_hash, load_intermediate as i32,
                                                 store_intermediate as i32, stream.raw())
    }, "compute_leaf_hashes")
}

pub fn compute_node_hashes(values: &DeviceSlice<u64>, results: &mut DeviceSlice<u64>, stream: &CudaStream)
                           -> CudaResult<()> {
    check(unsafe {
        sys::ba_poseidon2_gl_compute_node_hashes(values.as_ptr(), values.len(), results.as_mut_ptr(), results.len(),
                                                 stream.raw())
    }, "compute_node_hashes")
}

pub fn build_merkle_tree_leaves(values: &DeviceSlice<u64>, results: &mut DeviceSlice<u64>, log_rows_per_hash: u32,
                                load_intermediate: bool, store_intermediate: bool, stream: &CudaStream)
                                -> CudaResult<()> {
    check(unsafe {
        sys::ba_poseidon2_gl_build_merkle_tree_leaves(values.as_ptr(), values.len(), results.as_mut_ptr(),
                                                      results.len(), log_rows_per_hash, load_intermediate as i32,
                                                      store_intermediate as i32, stream.raw())
    }, "build_merkle_tree_leaves")
}

pub fn build_merkle_tree_nodes(values: &DeviceSlice<u64>, results: &mut DeviceSlice<u64>, layers_count: u32,
                               stream: &CudaStream) -> CudaResult<()> {
    check(unsafe {
        sys::ba_poseidon2_gl_build_merkle_tree_nodes(values.as_ptr(), values.len(), results.as_mut_ptr(),
                                                     results.len(), layers_count, stream.raw())
    }, "build_merkle_tree_nodes")
}

/// Same argument order as `GpuTreeHasher::build_merkle_tree`.
pub fn build_merkle_tree(values: &DeviceSlice<u64>, results: &mut DeviceSlice<u64>, log_rows_per_hash: u32,
                         stream: &CudaStream, layers_count: u32) -> CudaResult<()> {
    check(unsafe {
        sys::ba_poseidon2_gl_build_merkle_tree(values.as_ptr(), values.len(), results.as_mut_ptr(), results.len(),
                                               log_rows_per_hash, layers_count, stream.raw())
    }, "build_merkle_tree")
}
