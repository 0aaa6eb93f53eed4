//! Raw declarations of libboojum_amd.so, one for one with include/boojum_amd.h.
//!
//! Replaces, for the Goldilocks NTT / LDE path only, `cudart-sys` (crates/cudart-sys/src/
//! bindings.rs) and the `extern "C"` kernel entry points boojum-cuda binds
//! (crates/boojum-cuda/src/ntt.rs:15-56).  Every function returns 0 or an error code; status 1
//! (`BA_ERROR_INVALID_VALUE`) is where the reference panics on an `assert!`.
#![allow(non_camel_case_types)]
#![no_std]

use core::ffi::{c_char, c_int, c_void};

pub const BA_SUCCESS: c_int = 0;
pub const BA_ERROR_INVALID_VALUE: c_int = 1;
pub const BA_ERROR_OUT_OF_MEMORY: c_int = 2;
pub const BA_ERROR_NOT_INITIALIZED: c_int = 3;
pub const BA_OMEGA_LOG_ORDER: u32 = 30;
pub const BA_MAX_LOG_N: u32 = 28;
pub const BA_JOB_NTT: u32 = 0;
pub const BA_JOB_LDE: u32 = 1;
pub const BA_LDE_NATURAL_MONOMIALS: u32 = 1;
pub const BA_TUNING_LDE_COSETS: c_int = 2;
pub const BA_SERVE_INPUT_HOST: u32 = 0;
pub const BA_SERVE_INPUT_DEVICE: u32 = 1;
pub const BA_SERVE_INPUT_SYNTHETIC: u32 = 2;

pub type ba_stream_t = *mut c_void;
pub type ba_event_t = *mut c_void;
pub type ba_graph_exec_t = *mut c_void;
pub type ba_context_t = *mut c_void;
pub type ba_batch_plan_t = *mut c_void;
pub type ba_server_t = *mut c_void;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ba_job {
    pub kind: u32,
    pub log_n: u32,
    pub num_cols: u32,
    pub bitrev_inputs: u32,
    pub inverse: u32,
    pub log_extension_degree: u32,
    pub coset_index: u32,
    pub log_blowup: u32,
    pub inputs: *mut u64,
    pub outputs: *mut u64,
    pub monomials: *mut u64,
    pub stride_in: u64,
    pub stride_out: u64,
    pub out_coset_stride: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ba_tuned_plan {
    pub num_passes: u32,
    pub bits: [u32; 8],
    pub reps: [u32; 8],
    pub contig_waves: u32,
    pub tile_log: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ba_serve_request {
    pub log_n: u32,
    pub num_cols: u32,
    pub log_blowup: u32,
    pub input_kind: u32,
    pub evals: *const u64,
    pub in_stride: u64,
    pub seed: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ba_serve_result {
    pub status: i32,
    pub reserved: u32,
    pub outputs: *const u64,
    pub out_col_stride: u64,
    pub queue_ms: f64,
    pub latency_ms: f64,
    pub batch: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ba_serve_stats {
    pub completed: u64,
    pub failed: u64,
    pub batches: u64,
    pub output_elems: u64,
    pub p50_ms: f64,
    pub p99_ms: f64,
    pub mean_ms: f64,
    pub max_ms: f64,
    pub elapsed_s: f64,
    pub mean_batch_jobs: f64,
    pub budget_bytes: u64,
    pub peak_bytes: u64,
    pub in_use_bytes: u64,
}

extern "C" {
    // runtime subset (crates/cudart/src/{device,memory,stream,event,error}.rs)
    pub fn ba_get_device_count(count: *mut c_int) -> c_int;
    pub fn ba_set_device(device: c_int) -> c_int;
    pub fn ba_get_device(device: *mut c_int) -> c_int;
    pub fn ba_device_synchronize() -> c_int;
    pub fn ba_mem_get_info(free_bytes: *mut usize, total_bytes: *mut usize) -> c_int;
    pub fn ba_device_get_attribute(value: *mut c_int, attribute: c_int, device: c_int) -> c_int;
    pub fn ba_malloc(ptr: *mut *mut c_void, bytes: usize) -> c_int;
    pub fn ba_free(ptr: *mut c_void) -> c_int;
    pub fn ba_malloc_host(ptr: *mut *mut c_void, bytes: usize) -> c_int;
    pub fn ba_free_host(ptr: *mut c_void) -> c_int;
    pub fn ba_memcpy(dst: *mut c_void, src: *const c_void, bytes: usize, kind: c_int) -> c_int;
    pub fn ba_memcpy_async(dst: *mut c_void, src: *const c_void, bytes: usize, kind: c_int,
                           stream: ba_stream_t) -> c_int;
    pub fn ba_memset_async(dst: *mut c_void, value: c_int, bytes: usize, stream: ba_stream_t) -> c_int;
    pub fn ba_stream_create(stream: *mut ba_stream_t) -> c_int;
    pub fn ba_stream_destroy(stream: ba_stream_t) -> c_int;
    pub fn ba_stream_synchronize(stream: ba_stream_t) -> c_int;
    pub fn ba_stream_wait_event(stream: ba_stream_t, event: ba_event_t) -> c_int;
    pub fn ba_event_create(event: *mut ba_event_t) -> c_int;
    pub fn ba_event_destroy(event: ba_event_t) -> c_int;
    pub fn ba_event_record(event: ba_event_t, stream: ba_stream_t) -> c_int;
    pub fn ba_event_synchronize(event: ba_event_t) -> c_int;
    pub fn ba_event_query(event: ba_event_t) -> c_int;
    pub fn ba_event_elapsed_ms(ms: *mut f32, start: ba_event_t, end: ba_event_t) -> c_int;
    pub fn ba_get_last_error() -> c_int;
    pub fn ba_error_name(status: c_int) -> *const c_char;
    pub fn ba_stream_begin_capture(stream: ba_stream_t) -> c_int;
    pub fn ba_stream_end_capture(stream: ba_stream_t, exec: *mut ba_graph_exec_t) -> c_int;
    pub fn ba_graph_launch(exec: ba_graph_exec_t, stream: ba_stream_t) -> c_int;
    pub fn ba_graph_destroy(exec: ba_graph_exec_t) -> c_int;

    // boojum_cuda::context (crates/boojum-cuda/src/context.rs:171-296)
    pub fn ba_context_create(powers_of_w_coarse_log_count: u32, powers_of_g_coarse_log_count: u32,
                             ctx: *mut ba_context_t) -> c_int;
    pub fn ba_context_destroy(ctx: ba_context_t) -> c_int;

    // boojum_cuda::ntt (crates/boojum-cuda/src/ntt.rs:179-369)
    pub fn ba_batch_ntt_internal(inputs: *const u64, outputs: *mut u64, log_n: u32, num_ntts: u32,
                                 stride_between_input_arrays: u32, stride_between_output_arrays: u32,
                                 bitrev_inputs: c_int, inverse: c_int, log_extension_degree: u32,
                                 coset_index: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_batch_ntt_out_of_place(inputs_matrix: *const u64, inputs_len: usize, outputs_matrix: *mut u64,
                                     outputs_len: usize, log_n: u32, num_ntts: u32, inputs_offset: u32,
                                     outputs_offset: u32, stride_between_input_arrays: u32,
                                     stride_between_output_arrays: u32, bitrev_inputs: c_int, inverse: c_int,
                                     log_extension_degree: u32, coset_index: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_batch_ntt_in_place(inputs_matrix: *mut u64, inputs_len: usize, log_n: u32, num_ntts: u32,
                                 inputs_offset: u32, stride_between_input_arrays: u32, bitrev_inputs: c_int,
                                 inverse: c_int, log_extension_degree: u32, coset_index: u32,
                                 stream: ba_stream_t) -> c_int;
    pub fn ba_batch_ntt_ctx(ctx: ba_context_t, inputs: *const u64, outputs: *mut u64, log_n: u32, num_ntts: u32,
                            stride_in: u64, stride_out: u64, bitrev_inputs: c_int, inverse: c_int,
                            log_extension_degree: u32, coset_index: u32, stream: ba_stream_t) -> c_int;

    // boojum_cuda::ops_complex (crates/boojum-cuda/src/ops_complex.rs:31-331)
    pub fn ba_bit_reverse(src: *const u64, src_stride: u64, dst: *mut u64, dst_stride: u64, log_rows: u32,
                          cols: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_bit_reverse_in_place(values: *mut u64, stride: u64, log_rows: u32, cols: u32,
                                   stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_of_w(log_degree: u32, offset: u32, inverse: c_int, bit_reverse: c_int, result: *mut u64,
                              count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_of_g(log_degree: u32, offset: u32, inverse: c_int, bit_reverse: c_int, result: *mut u64,
                              count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_by_val_bf(base: u64, offset: u32, bit_reverse: c_int, result: *mut u64, count: u32,
                                   stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_by_ref_bf(base: *const u64, offset: u32, bit_reverse: c_int, result: *mut u64,
                                   count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_by_val_ef(base_c0: u64, base_c1: u64, offset: u32, bit_reverse: c_int,
                                   result: *mut u64, count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_get_powers_by_ref_ef(base: *const u64, offset: u32, bit_reverse: c_int, result: *mut u64,
                                   count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_omega_shift(src: *const u64, log_degree: u32, offset: u32, inverse: c_int, shift: u32,
                          dst: *mut u64, count: u32, stream: ba_stream_t) -> c_int;
    pub fn ba_omega_shift_in_place(values: *mut u64, log_degree: u32, offset: u32, inverse: c_int, shift: u32,
                                   count: u32, stream: ba_stream_t) -> c_int;

    // boojum_cuda::barycentric (crates/boojum-cuda/src/barycentric.rs:83-300)
    pub fn ba_barycentric_precompute_lagrange_coeffs(ext: c_int, x: *const u64, common_factor: *mut u64,
                                                     coset: u64, lagrange_coeffs: *mut u64, count: u32,
                                                     bit_reverse: c_int, stream: ba_stream_t) -> c_int;
    pub fn ba_barycentric_batch_eval_temp_elems(kind: c_int, count: u32, num_polys: u32, elems: *mut u64) -> c_int;
    pub fn ba_barycentric_batch_eval(kind: c_int, ys: *const u64, count: u32, num_polys: u32,
                                     lagrange_coeffs: *const u64, temp: *mut u64, temp_elems: u64,
                                     evals: *mut u64, stream: ba_stream_t) -> c_int;
    // no reference counterpart: batch_eval for callers that cannot allocate (per-GPU array inside)
    pub fn ba_barycentric_batch_eval_noalloc(kind: c_int, ys: *const u64, count: u32, num_polys: u32,
                                             lagrange_coeffs: *const u64, evals: *mut u64,
                                             stream: ba_stream_t) -> c_int;

    // shivini's LDE composition in one call (crates/shivini/src/data_structures/storage.rs:154-260)
    pub fn ba_lde_batch(ctx: ba_context_t, evals: *mut u64, in_stride: u64, monomials: *mut u64, out: *mut u64,
                        out_col_stride: u64, out_coset_stride: u64, log_n: u32, num_cols: u32, log_blowup: u32,
                        stream: ba_stream_t) -> c_int;
    pub fn ba_lde_batch_ex(ctx: ba_context_t, evals: *mut u64, in_stride: u64, monomials: *mut u64, out: *mut u64,
                           out_col_stride: u64, out_coset_stride: u64, log_n: u32, num_cols: u32, log_blowup: u32,
                           flags: u32, stream: ba_stream_t) -> c_int;

    // uneven batches, autotuning, serve mode (no reference counterpart)
    pub fn ba_batch_plan_create(ctx: ba_context_t, jobs: *const ba_job, num_jobs: u32,
                                plan: *mut ba_batch_plan_t) -> c_int;
    pub fn ba_batch_plan_launch(plan: ba_batch_plan_t, stream: ba_stream_t) -> c_int;
    pub fn ba_batch_plan_stats(plan: ba_batch_plan_t, launches: *mut u32, levels: *mut u32,
                               workgroups: *mut u64) -> c_int;
    pub fn ba_batch_plan_destroy(plan: ba_batch_plan_t) -> c_int;
    pub fn ba_autotune(ctx: ba_context_t, log_n_min: u32, log_n_max: u32, elems_per_trial: u64,
                       stream: ba_stream_t) -> c_int;
    pub fn ba_tuning_get(ctx: ba_context_t, log_n: u32, bitrev_inputs: c_int, plan: *mut ba_tuned_plan) -> c_int;
    pub fn ba_tuning_set(ctx: ba_context_t, log_n: u32, bitrev_inputs: c_int, plan: *const ba_tuned_plan) -> c_int;
    pub fn ba_tuning_candidates(log_n: u32, plans: *mut ba_tuned_plan, capacity: u32, count: *mut u32) -> c_int;
    pub fn ba_server_create(ctx: ba_context_t, memory_budget_bytes: usize, max_inflight_batches: u32,
                            server: *mut ba_server_t) -> c_int;
    pub fn ba_server_set_max_bypass(server: ba_server_t, max_bypass: u32) -> c_int;
    pub fn ba_server_submit(server: ba_server_t, request: *const ba_serve_request, ticket: *mut u64) -> c_int;
    pub fn ba_server_wait(server: ba_server_t, ticket: u64, result: *mut ba_serve_result) -> c_int;
    pub fn ba_server_poll(server: ba_server_t, ticket: u64, done: *mut c_int, result: *mut ba_serve_result) -> c_int;
    pub fn ba_server_release(server: ba_server_t, ticket: u64) -> c_int;
    pub fn ba_server_stats(server: ba_server_t, stats: *mut ba_serve_stats) -> c_int;
    pub fn ba_server_destroy(server: ba_server_t) -> c_int;

    // diagnostic: the VALU rate of the passes' butterfly arithmetic in registers
    pub fn ba_diag_butterfly_rate(input: *const u64, output: *mut u64, blocks: u32, iters: u32,
                                  stream: ba_stream_t) -> c_int;

    // Poseidon2 Merkle hashing (crates/boojum-cuda/src/poseidon2.rs:165-316): slices as
    // (pointer, length in u64 words); the context's round constants must be installed first
    pub fn ba_poseidon2_gl_set_round_constants(ctx: ba_context_t, round_constants: *const u64) -> c_int;
    pub fn ba_poseidon2_gl_compute_leaf_hashes(values: *const u64, values_len: usize, results: *mut u64,
                                               results_len: usize, log_rows_per_hash: u32,
                                               load_intermediate: c_int, store_intermediate: c_int,
                                               stream: ba_stream_t) -> c_int;
    pub fn ba_poseidon2_gl_compute_node_hashes(values: *const u64, values_len: usize, results: *mut u64,
                                               results_len: usize, stream: ba_stream_t) -> c_int;
    pub fn ba_poseidon2_gl_build_merkle_tree_leaves(values: *const u64, values_len: usize, results: *mut u64,
                                                    results_len: usize, log_rows_per_hash: u32,
                                                    load_intermediate: c_int, store_intermediate: c_int,
                                                    stream: ba_stream_t) -> c_int;
    pub fn ba_poseidon2_gl_build_merkle_tree_nodes(values: *const u64, values_len: usize, results: *mut u64,
                                                   results_len: usize, layers_count: u32,
                                                   stream: ba_stream_t) -> c_int;
    pub fn ba_poseidon2_gl_build_merkle_tree(values: *const u64, values_len: usize, results: *mut u64,
                                             results_len: usize, log_rows_per_hash: u32, layers_count: u32,
                                             stream: ba_stream_t) -> c_int;

    // launch tracing and helpers
    pub fn ba_profile_begin(events: *mut ba_event_t, capacity: u32) -> c_int;
    pub fn ba_profile_end(launches: *mut u32) -> c_int;
    pub fn ba_profile_launch_name(index: u32) -> *const c_char;
    pub fn ba_fill_uniform(dst: *mut u64, stride: u64, rows: u64, cols: u32, seed: u64,
                           stream: ba_stream_t) -> c_int;
    pub fn ba_field_op(op: c_int, a: *const u64, b: *const u64, out: *mut u64, n: usize,
                       stream: ba_stream_t) -> c_int;
}

#[cfg(test)]
mod tests {
    // the C header's sizes (checked against a C compiler by tests/test_abi.py)
    #[test]
    fn struct_sizes() {
        assert_eq!(core::mem::size_of::<super::ba_job>(), 80);
        assert_eq!(core::mem::size_of::<super::ba_tuned_plan>(), 68);
        assert_eq!(core::mem::size_of::<super::ba_serve_request>(), 40);
        assert_eq!(core::mem::size_of::<super::ba_serve_result>(), 48);
        assert_eq!(core::mem::size_of::<super::ba_serve_stats>(), 104);
    }
}
