/*
 * boojum_amd.h -- C-ABI of the MI355X-native Goldilocks NTT / coset-LDE library
 * (libboojum_amd.so, built from zksync-crypto-gpu_amd/csrc for gfx950).
 *
 * It replaces, for the NTT/LDE path only:
 *   - era_cudart / cudart-sys (crates/cudart, crates/cudart-sys): the runtime subset the
 *     path rides on (streams, events, memory, graphs), as thin HIP wrappers;
 *   - boojum_cuda::context (crates/boojum-cuda/src/context.rs): twiddle / coset tables;
 *   - boojum_cuda::ntt (crates/boojum-cuda/src/ntt.rs): the batch NTT entry points;
 *   - boojum_cuda::ops_complex::bit_reverse* (crates/boojum-cuda/src/ops_complex.rs:262-331).
 *
 * Conventions
 *   - Every function returns an int status: 0 (BA_SUCCESS) or a HIP error code; argument
 *     preconditions that the reference enforces with assert! return BA_ERROR_INVALID_VALUE
 *     (hipErrorInvalidValue) instead of panicking.  Nothing here aborts the process.
 *   - Field elements are uint64_t.  Inputs may hold any u64 (read mod p); outputs are
 *     canonical (< p).  Matrices are column-major: column k starts at ptr + k * stride.
 *   - Device pointers are plain pointers; streams/events/graphs are opaque handles (a null
 *     stream is the default stream).  All compute calls are asynchronous on their stream and
 *     borrow their buffers: the caller keeps them alive until the stream is synchronised.
 */
#ifndef BOOJUM_AMD_H
#define BOOJUM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BA_SUCCESS 0
#define BA_ERROR_INVALID_VALUE 1         /* hipErrorInvalidValue */
#define BA_ERROR_OUT_OF_MEMORY 2         /* hipErrorOutOfMemory */
#define BA_ERROR_NOT_INITIALIZED 3       /* hipErrorNotInitialized: no live context */

typedef void *ba_stream_t;     /* hipStream_t */
typedef void *ba_event_t;      /* hipEvent_t */
typedef void *ba_graph_exec_t; /* hipGraphExec_t */
typedef struct ba_context *ba_context_t;

/* ---------------------------------------------------------------- runtime subset
 * Replaces the cudart calls the path uses (crates/cudart/src/{device,memory,stream,event,
 * error}.rs); each is a 1:1 HIP call with its error code returned as int. */
int ba_get_device_count(int *count);
int ba_set_device(int device);
int ba_get_device(int *device);
int ba_device_synchronize(void);
int ba_mem_get_info(size_t *free_bytes, size_t *total_bytes);
/* hipDeviceAttribute_t value; e.g. 16 = multiprocessor count */
int ba_device_get_attribute(int *value, int attribute, int device);
int ba_malloc(void **ptr, size_t bytes);
int ba_free(void *ptr);
int ba_malloc_host(void **ptr, size_t bytes); /* pinned */
int ba_free_host(void *ptr);
/* kind: 0 host->host, 1 host->device, 2 device->host, 3 device->device, 4 default */
int ba_memcpy(void *dst, const void *src, size_t bytes, int kind);
int ba_memcpy_async(void *dst, const void *src, size_t bytes, int kind, ba_stream_t stream);
int ba_memset_async(void *dst, int value, size_t bytes, ba_stream_t stream);
int ba_stream_create(ba_stream_t *stream); /* non-blocking, default priority */
int ba_stream_destroy(ba_stream_t stream);
int ba_stream_synchronize(ba_stream_t stream);
int ba_stream_wait_event(ba_stream_t stream, ba_event_t event);
int ba_event_create(ba_event_t *event);
int ba_event_destroy(ba_event_t event);
int ba_event_record(ba_event_t event, ba_stream_t stream);
int ba_event_synchronize(ba_event_t event);
int ba_event_query(ba_event_t event); /* 0 done, hipErrorNotReady (600) pending */
int ba_event_elapsed_ms(float *ms, ba_event_t start, ba_event_t end);
int ba_get_last_error(void);
const char *ba_error_name(int status);
int ba_stream_begin_capture(ba_stream_t stream);
int ba_stream_end_capture(ba_stream_t stream, ba_graph_exec_t *exec);
int ba_graph_launch(ba_graph_exec_t exec, ba_stream_t stream);
int ba_graph_destroy(ba_graph_exec_t exec);

/* ---------------------------------------------------------------- context
 * boojum_cuda::context::Context::create / destroy (crates/boojum-cuda/src/context.rs:171-296).
 * Builds the power tables of omega_{2^30}, g = 7 and g^-1 split into fine/coarse halves
 * (coarse sizes 2^w_coarse_log and 2^g_coarse_log, each <= 30) plus 2^-k for k <= 30, and
 * keeps them resident in HBM.  The most recently created live context of the current device
 * is the one the context-free entry points below use, as the reference's process-global
 * tables are (context.rs:55-60). */
int ba_context_create(uint32_t powers_of_w_coarse_log_count, uint32_t powers_of_g_coarse_log_count,
                      ba_context_t *ctx);
int ba_context_destroy(ba_context_t ctx);
#define BA_OMEGA_LOG_ORDER 30u
/* Largest column handled: 2^28 elements (the reference's plans end at 2^24, ntt.rs:74-169). */
#define BA_MAX_LOG_N 28u

/* ---------------------------------------------------------------- NTT
 * boojum_cuda::ntt::batch_ntt_internal (crates/boojum-cuda/src/ntt.rs:179-291).
 * num_ntts columns of 2^log_n elements; input column k at inputs + k*stride_in, output at
 * outputs + k*stride_out (inputs == outputs: in place).
 *   bitrev_inputs=0, inverse=0: out[bitrev(k)] = sum_j s^j x_j w^(jk)       (natural -> bitrev)
 *   bitrev_inputs=1, inverse=0: input bit-reversed, output natural
 *   bitrev_inputs=0, inverse=1: out[bitrev(j)] = n^-1 s^-j sum_k X_k w^(-jk)
 *   bitrev_inputs=1, inverse=1: input bit-reversed, output natural
 * with w = omega_{2^log_n} and s = 1 if log_extension_degree == 0 else
 * 7 * omega_{2^(log_n + log_extension_degree)}^coset_index.
 * Preconditions (ntt.rs:192-200): 1 <= log_n <= BA_MAX_LOG_N, log_n + log_extension_degree
 * <= 30, 2^log_n <= strides, coset_index < 2^log_extension_degree when nonzero. */
int ba_batch_ntt_internal(const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                          uint32_t num_ntts, uint32_t stride_between_input_arrays,
                          uint32_t stride_between_output_arrays, int bitrev_inputs, int inverse,
                          uint32_t log_extension_degree, uint32_t coset_index, ba_stream_t stream);

/* boojum_cuda::ntt::batch_ntt_out_of_place (ntt.rs:294-335): DeviceSlices passed as
 * (pointer, length in elements); checks offset + n <= stride, lengths >= num_ntts * stride
 * and that the (offset) pointers differ. */
int ba_batch_ntt_out_of_place(const uint64_t *inputs_matrix, size_t inputs_len,
                              uint64_t *outputs_matrix, size_t outputs_len, uint32_t log_n,
                              uint32_t num_ntts, uint32_t inputs_offset, uint32_t outputs_offset,
                              uint32_t stride_between_input_arrays,
                              uint32_t stride_between_output_arrays, int bitrev_inputs,
                              int inverse, uint32_t log_extension_degree, uint32_t coset_index,
                              ba_stream_t stream);

/* boojum_cuda::ntt::batch_ntt_in_place (ntt.rs:338-369). */
int ba_batch_ntt_in_place(uint64_t *inputs_matrix, size_t inputs_len, uint32_t log_n,
                          uint32_t num_ntts, uint32_t inputs_offset,
                          uint32_t stride_between_input_arrays, int bitrev_inputs, int inverse,
                          uint32_t log_extension_degree, uint32_t coset_index, ba_stream_t stream);

/* Same as ba_batch_ntt_internal with an explicit context (several may be live). */
int ba_batch_ntt_ctx(ba_context_t ctx, const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                     uint32_t num_ntts, uint64_t stride_in, uint64_t stride_out, int bitrev_inputs,
                     int inverse, uint32_t log_extension_degree, uint32_t coset_index,
                     ba_stream_t stream);

/* ---------------------------------------------------------------- bit reverse
 * boojum_cuda::ops_complex::bit_reverse / bit_reverse_in_place (ops_complex.rs:262-331):
 * dst[k][bitrev(i)] = src[k][i] for cols columns of 2^log_rows rows (src == dst: in place). */
int ba_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst, uint64_t dst_stride,
                   uint32_t log_rows, uint32_t cols, ba_stream_t stream);
int ba_bit_reverse_in_place(uint64_t *values, uint64_t stride, uint32_t log_rows, uint32_t cols,
                            ba_stream_t stream);

/* ---------------------------------------------------------------- coset LDE
 * shivini's trace-commitment LDE over num_cols columns (crates/shivini/src/data_structures/
 * storage.rs:154-260 with primitives/ntt.rs:10-42): evaluations (natural order) are
 * interpolated to monomials, then for each coset c < 2^log_blowup the coset evaluations
 * f(7 * omega_{n 2^log_blowup}^bitrev(c) * omega_n^bitrev(i)) are written bit-reversed to
 * out + k*out_col_stride + c*out_coset_stride.  With out_col_stride = n << log_blowup and
 * out_coset_stride = n a column holds its whole LDE domain in bit-reversed order; with
 * out_col_stride = n and out_coset_stride = num_cols * n the cosets are separate matrices, as
 * shivini stores them.  The monomials, in bit-reversed order, go to `monomials` (stride n)
 * when it is non-null, else they replace the evaluations in place (as into_monomials consumes
 * its storage, storage.rs:155).  Preconditions: 1 <= log_n, 1 <= log_blowup,
 * log_n + log_blowup <= 30, strides >= n, out distinct from evals / monomials, the context's
 * device current.  Cosets after the first run on auxiliary streams of the context, forked from
 * `stream` after the inverse NTT and joined back into it before the call returns (a graph
 * captured from `stream` holds the same fork and join; under ba_profile_begin, or with
 * BOOJUM_AMD_LDE_SERIAL=1 in the environment, every launch stays on `stream`, in order); calls
 * on one context are serialised
 * while they enqueue, and no call on the context may run in another thread while a stream
 * capture holding one of its LDEs is still open. */
int ba_lde_batch(ba_context_t ctx, uint64_t *evals, uint64_t in_stride, uint64_t *monomials,
                 uint64_t *out, uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                 uint32_t num_cols, uint32_t log_blowup, ba_stream_t stream);
/* ba_lde_batch with options.  BA_LDE_NATURAL_MONOMIALS: the monomials are left in natural order
 * (coefficient i at position i), as shivini's into_monomials returns them after its bit-reverse
 * epilogue (crates/shivini/src/data_structures/storage.rs:155, primitives/ntt.rs:572-585): one
 * in-place bit reverse of the monomials on `stream` after every coset has read them.  Any other
 * flag bit is an invalid value. */
#define BA_LDE_NATURAL_MONOMIALS 1u
int ba_lde_batch_ex(ba_context_t ctx, uint64_t *evals, uint64_t in_stride, uint64_t *monomials,
                    uint64_t *out, uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                    uint32_t num_cols, uint32_t log_blowup, uint32_t flags, ba_stream_t stream);

/* ---------------------------------------------------------------- powers, omega shifts
 * boojum_cuda::ops_complex (crates/boojum-cuda/src/ops_complex.rs:31-260): vectors of powers
 * read from the current device's context tables, same semantics and u32 index wrapping as the
 * reference kernels (crates/boojum-cuda/native/ops_complex.cu:8-86):
 *   get_powers_of_w: result[i] = omega_{2^30}^(+-(j << (30 - log_degree))), j = i + offset, or its
 *     low log_degree bits reversed when bit_reverse (omega_{2^log_degree}^j for j < 2^log_degree);
 *   get_powers_of_g: 7^(the same exponent mod 2^30), 7^-1 when inverse;
 *   get_powers_by_val/ref_bf: result[i] = base^((bit_reverse ? brev32(i) : i) + offset);
 *   get_powers_by_val/ref_ef: the same in F_p[u]/(u^2 - 7), stored as two planes:
 *     result[i] = c0, result[count + i] = c1 (the reference's extension_field layout);
 *   omega_shift: dst[i] = src[i] * omega_{2^30}^(+-((shift * (i + offset)) << (30 - log_degree)));
 *     dst == src is the in-place form (ops_complex.rs:246-260); other overlaps are rejected.
 * log_degree <= 30. */
int ba_get_powers_of_w(uint32_t log_degree, uint32_t offset, int inverse, int bit_reverse,
                       uint64_t *result, uint32_t count, ba_stream_t stream);
int ba_get_powers_of_g(uint32_t log_degree, uint32_t offset, int inverse, int bit_reverse,
                       uint64_t *result, uint32_t count, ba_stream_t stream);
int ba_get_powers_by_val_bf(uint64_t base, uint32_t offset, int bit_reverse, uint64_t *result,
                            uint32_t count, ba_stream_t stream);
int ba_get_powers_by_ref_bf(const uint64_t *base, uint32_t offset, int bit_reverse,
                            uint64_t *result, uint32_t count, ba_stream_t stream);
int ba_get_powers_by_val_ef(uint64_t base_c0, uint64_t base_c1, uint32_t offset, int bit_reverse,
                            uint64_t *result, uint32_t count, ba_stream_t stream);
int ba_get_powers_by_ref_ef(const uint64_t *base, uint32_t offset, int bit_reverse,
                            uint64_t *result, uint32_t count, ba_stream_t stream);
int ba_omega_shift(const uint64_t *src, uint32_t log_degree, uint32_t offset, int inverse,
                   uint32_t shift, uint64_t *dst, uint32_t count, ba_stream_t stream);
int ba_omega_shift_in_place(uint64_t *values, uint32_t log_degree, uint32_t offset, int inverse,
                            uint32_t shift, uint32_t count, ba_stream_t stream);

/* ---------------------------------------------------------------- barycentric evaluation
 * boojum_cuda::barycentric (crates/boojum-cuda/src/barycentric.rs:83-300): evaluations y_i of
 * polynomials on the coset domain {coset * w^i : i < count} (w = omega_count, count a power of
 * two) are evaluated at a point X as f(X) = sum_i L_i y_i, with
 * L_i = common / (X w^-i - coset) and common = coset (X^count - coset^count) / (count coset^count).
 * ext = 0: X, common_factor (1 element each) and the count coefficients are base-field;
 * ext = 1: X and common_factor are (c0, c1) in F_p[u]/(u^2 - 7) and the coefficients are two
 * planes c0[count], c1[count].  bit_reverse stores the coefficients in bit-reversed order (for
 * evaluations stored bit-reversed, as the LDE leaves them).  x and common_factor are device
 * pointers (DeviceVariable in the reference); the tables are the current device's context.
 * batch_eval kind: 0 base-field ys at a base-field X, 1 base-field ys at an extension X,
 * 2 extension ys (column k at ys + 2 k count: planes c0, c1) at an extension X.  ys columns are
 * count apart (count for kind 2: planes).  evals: num_polys base elements (kind 0) or
 * num_polys (c0, c1) pairs.  temp: device scratch of ba_barycentric_batch_eval_temp_elems
 * elements (the reference's temp_storage arguments, barycentric.rs:193-209). */
int ba_barycentric_precompute_lagrange_coeffs(int ext, const uint64_t *x, uint64_t *common_factor,
                                              uint64_t coset, uint64_t *lagrange_coeffs,
                                              uint32_t count, int bit_reverse, ba_stream_t stream);
int ba_barycentric_batch_eval_temp_elems(int kind, uint32_t count, uint32_t num_polys,
                                         uint64_t *elems);
int ba_barycentric_batch_eval(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                              const uint64_t *lagrange_coeffs, uint64_t *temp, uint64_t temp_elems,
                              uint64_t *evals, ba_stream_t stream);
/* No counterpart in the reference.  ba_barycentric_batch_eval for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses that forbid it): the
 * partial sums go to a 4 MiB array the library holds on each GPU until the process ends, so the
 * call needs no temp argument and allocates nothing.  Requests whose
 * ba_barycentric_batch_eval_temp_elems exceeds 2^19 elements are an invalid value.  Every call on
 * a GPU uses that GPU's one array: calls that overlap on one GPU (different streams, different
 * threads) race. */
int ba_barycentric_batch_eval_noalloc(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                                      const uint64_t *lagrange_coeffs, uint64_t *evals, ba_stream_t stream);

/* ---------------------------------------------------------------- uneven batches
 * No reference counterpart (boojum-cuda runs one transform per call, ntt.rs:179-291): a list of
 * independent jobs -- batch NTTs of any size, direction and coset, and coset LDEs -- planned
 * once into a few grouped launches (one per dependency level and pass shape) and launched any
 * number of times, e.g. inside a captured hipGraph.  Jobs must not alias each other's buffers.
 * The plan holds device copies of the job descriptors; buffers are not owned. */
#define BA_JOB_NTT 0u
#define BA_JOB_LDE 1u
typedef struct {
  uint32_t kind;                 /* BA_JOB_NTT or BA_JOB_LDE */
  uint32_t log_n;
  uint32_t num_cols;
  uint32_t bitrev_inputs;        /* NTT: as ba_batch_ntt_internal */
  uint32_t inverse;              /* NTT */
  uint32_t log_extension_degree; /* NTT */
  uint32_t coset_index;          /* NTT */
  uint32_t log_blowup;           /* LDE: as ba_lde_batch */
  uint64_t *inputs;              /* NTT input / LDE evaluations */
  uint64_t *outputs;             /* NTT output / LDE coset evaluations */
  uint64_t *monomials;           /* LDE only, may be NULL */
  uint64_t stride_in;            /* column stride of inputs */
  uint64_t stride_out;           /* column stride of outputs */
  uint64_t out_coset_stride;     /* LDE only */
} ba_job;
typedef struct ba_batch_plan *ba_batch_plan_t;
int ba_batch_plan_create(ba_context_t ctx, const ba_job *jobs, uint32_t num_jobs,
                         ba_batch_plan_t *plan);
int ba_batch_plan_launch(ba_batch_plan_t plan, ba_stream_t stream);
int ba_batch_plan_stats(ba_batch_plan_t plan, uint32_t *launches, uint32_t *levels,
                        uint64_t *workgroups);
int ba_batch_plan_destroy(ba_batch_plan_t plan);

/* ---------------------------------------------------------------- autotuning
 * No reference counterpart (boojum-cuda's pass shapes are fixed tables, ntt.rs:57-176).  A size
 * class is (log_n, input order); its plan is the split of the log_n index bits into passes --
 * bits[0] for the contiguous pass over the lowest bits, then the strided passes upwards -- and,
 * per pass in the same order, the columns each workgroup repeats over (reps, 1..8; 0 = built-in
 * choice), and the waves per SIMD of a 9..12-bit contiguous pass (contig_waves: 3 = the variant
 * that loads the next tile while transforming, 4 = the 128-register variant without that
 * prefetch, 0 = built-in choice: 3 below 2048 workgroups, else 4), and the tile shape of plain
 * batch NTTs (tile_log: 10 = the latency-shaped passes, 1024-element tiles of 4 elements per
 * thread in radix-4 rounds, for every batch size, log_n 4..18 only; 12 = the 4096-element tiles
 * of the split above for every batch size; 0 = built-in choice: 1024-element tiles for batches of
 * at most 2^20 elements, else 4096).  LDE coset NTTs always run the 4096-element tiles.  Valid splits: log_n <= 12
 * one pass of log_n bits; log_n > 12 a contiguous pass of 9..12 bits and strided passes of 4..10
 * bits (at most 8 passes).
 *
 * ba_autotune times every valid split, then each pass's reps, then contig_waves, then (log_n 4..18)
 * the 1024-element tiles against the best of those, for plain NTTs of both input orders and for
 * the LDE's coset NTTs (BA_TUNING_LDE_COSETS below), for log_n in [log_n_min,
 * log_n_max] and both input orders, on batches of max(elems_per_trial, 2^log_n) elements of
 * scratch memory it allocates and frees, and keeps the fastest on `ctx`; every later transform
 * (NTT, LDE, batch plan) on `ctx` uses it.  Results are bit-identical for every plan.  Tuning
 * state is per context and must not change while other threads launch with that context. */
typedef struct {
  uint32_t num_passes;
  uint32_t bits[8];
  uint32_t reps[8];
  uint32_t contig_waves; /* 0, 3 or 4 */
  uint32_t tile_log;     /* 0, 10 or 12 */
} ba_tuned_plan;
int ba_autotune(ba_context_t ctx, uint32_t log_n_min, uint32_t log_n_max, uint64_t elems_per_trial,
                ba_stream_t stream);
/* bitrev_inputs = BA_TUNING_LDE_COSETS addresses the plan of the LDE's coset NTTs (natural ->
 * bit-reversed forward NTTs that gather the bit-reversed monomials and scale by the coset in their
 * first pass), which ba_autotune times on that transform: its strided passes are at most 8 bits
 * and its tile_log is 0.  0 and 1 address plain batch NTTs of natural and bit-reversed input
 * (the LDE's inverse NTT follows 0). */
#define BA_TUNING_LDE_COSETS 2
/* the plan in use for (log_n, bitrev_inputs): tuned if set, else the built-in one */
int ba_tuning_get(ba_context_t ctx, uint32_t log_n, int bitrev_inputs, ba_tuned_plan *plan);
/* sets (plan != NULL, validated) or clears (plan == NULL) the tuned plan of a size class */
int ba_tuning_set(ba_context_t ctx, uint32_t log_n, int bitrev_inputs, const ba_tuned_plan *plan);
/* the splits ba_autotune tries for log_n (host only, no device needed): up to `capacity` plans
 * written to `plans` (bits set, reps 0); *count receives the total number of candidates */
int ba_tuning_candidates(uint32_t log_n, ba_tuned_plan *plans, uint32_t capacity, uint32_t *count);

/* ---------------------------------------------------------------- serve mode
 * No reference counterpart (shivini runs one prover's LDEs in program order,
 * crates/shivini/src/data_structures/storage.rs:154-260): a request loop for coset-LDE jobs.
 * Clients submit from any thread; the server's scheduler thread admits requests in arrival
 * order while its device arena (one allocation of memory_budget_bytes) has room for the
 * request's working buffer (2^log_n x num_cols elements) and outputs (2^(log_n+log_blowup) x
 * num_cols) -- while the oldest queued request waits for memory, later requests that fit may pass
 * it, at most max_bypass of them (default 32, ba_server_set_max_bypass; 0 = strict arrival
 * order) --, merges the requests admitted since its last launch into batch plans, one per output
 * size class (below 2^26 elements, below 2^30, from 2^30 up; the oldest admitted request's class
 * first), and enqueues each on a free slot stream (max_inflight_batches slots), so batches overlap
 * on the device and a small request never waits for a large one merged with it.
 * Results are the same as ba_lde_batch on the same evaluations: column k, coset c of the
 * output is at outputs + k * out_col_stride + c * 2^log_n (bit-reversed within the coset, the
 * cosets in shivini's order).  Outputs stay resident until ba_server_release.  A request
 * larger than the whole budget is refused by ba_server_submit with BA_ERROR_OUT_OF_MEMORY. */
#define BA_SERVE_INPUT_HOST 0u      /* evals in host memory, copied in; valid until completion */
#define BA_SERVE_INPUT_DEVICE 1u    /* evals in device memory, read (not modified) until completion */
#define BA_SERVE_INPUT_SYNTHETIC 2u /* no evals: ba_fill_uniform(.., seed) on the device */
typedef struct {
  uint32_t log_n;
  uint32_t num_cols;
  uint32_t log_blowup;   /* 1.. ; log_n + log_blowup <= BA_OMEGA_LOG_ORDER */
  uint32_t input_kind;   /* BA_SERVE_INPUT_* */
  const uint64_t *evals; /* column-major, column k at evals + k * in_stride */
  uint64_t in_stride;    /* 0 = 2^log_n */
  uint64_t seed;         /* BA_SERVE_INPUT_SYNTHETIC */
} ba_serve_request;
typedef struct {
  int32_t status;            /* 0, or the error that failed the request */
  uint32_t reserved;
  const uint64_t *outputs;   /* device memory owned by the server until released */
  uint64_t out_col_stride;   /* 2^(log_n + log_blowup) */
  double queue_ms;           /* submit -> admitted (memory granted) */
  double latency_ms;         /* submit -> completion seen by the scheduler */
  uint64_t batch;            /* the batch the request ran in */
} ba_serve_result;
typedef struct {
  uint64_t completed, failed, batches, output_elems;
  double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests */
  double elapsed_s;                       /* first submit -> last completion */
  double mean_batch_jobs;
  uint64_t budget_bytes, peak_bytes, in_use_bytes;
} ba_serve_stats;
typedef struct ba_server *ba_server_t;
int ba_server_create(ba_context_t ctx, size_t memory_budget_bytes, uint32_t max_inflight_batches,
                     ba_server_t *server);
/* how many later requests may be admitted ahead of a queued request waiting for memory */
int ba_server_set_max_bypass(ba_server_t server, uint32_t max_bypass);
int ba_server_submit(ba_server_t server, const ba_serve_request *request, uint64_t *ticket);
/* blocks until the request completed or failed */
int ba_server_wait(ba_server_t server, uint64_t ticket, ba_serve_result *result);
/* *done = 1 (and *result filled, if given) once the request completed or failed */
int ba_server_poll(ba_server_t server, uint64_t ticket, int *done, ba_serve_result *result);
/* returns a finished request's outputs to the arena; the ticket is then forgotten */
int ba_server_release(ba_server_t server, uint64_t ticket);
int ba_server_stats(ba_server_t server, ba_serve_stats *stats);
/* completes every submitted request, then frees the arena and the slots; no other call on the
 * same server may run concurrently with it or follow it */
int ba_server_destroy(ba_server_t server);

/* ---------------------------------------------------------------- Poseidon2 Merkle hashing
 * boojum_cuda::poseidon2::GpuTreeHasher for the Goldilocks Poseidon2 hasher
 * (crates/boojum-cuda/src/poseidon2.rs:165-316; kernels
 * crates/boojum-cuda/native/poseidon2/gl/poseidon2_gl_st.cu:24-88): width 12, rate 8, capacity
 * 4, the paper's permutation (eprint 2023/323) with 4 + 22 + 4 rounds.  Slices are (pointer,
 * length in u64 words) with the Rust functions' lengths, layouts and preconditions (a violated
 * assert! returns BA_ERROR_INVALID_VALUE before anything is launched).  Digests are 4 planes:
 * word k of digest g at results[g + k * count].  The calls use the current device's context and
 * return BA_ERROR_NOT_INITIALIZED until its round constants are installed. */
/* installs boojum's 30 x 12 round constants (poseidon_goldilocks_params
 * ALL_ROUND_CONSTANTS_AS_FIELD_ELEMENTS, row-major by round; read mod p) into ctx; no reference
 * counterpart: the reference compiles them into its kernels (build/poseidon2_gl.rs).  The copy is
 * synchronous; replacing constants while hashing work on the context is in flight races with it */
int ba_poseidon2_gl_set_round_constants(ba_context_t ctx, const uint64_t *round_constants);
/* GpuTreeHasher::compute_leaf_hashes (poseidon2.rs:180-229) */
int ba_poseidon2_gl_compute_leaf_hashes(const uint64_t *values, size_t values_len, uint64_t *results,
                                        size_t results_len, uint32_t log_rows_per_hash, int load_intermediate,
                                        int store_intermediate, ba_stream_t stream);
/* GpuTreeHasher::compute_node_hashes (poseidon2.rs:231-251) */
int ba_poseidon2_gl_compute_node_hashes(const uint64_t *values, size_t values_len, uint64_t *results,
                                        size_t results_len, ba_stream_t stream);
/* GpuTreeHasher::build_merkle_tree_leaves (poseidon2.rs:253-274) */
int ba_poseidon2_gl_build_merkle_tree_leaves(const uint64_t *values, size_t values_len, uint64_t *results,
                                             size_t results_len, uint32_t log_rows_per_hash, int load_intermediate,
                                             int store_intermediate, ba_stream_t stream);
/* GpuTreeHasher::build_merkle_tree_nodes (poseidon2.rs:276-295) */
int ba_poseidon2_gl_build_merkle_tree_nodes(const uint64_t *values, size_t values_len, uint64_t *results,
                                            size_t results_len, uint32_t layers_count, ba_stream_t stream);
/* GpuTreeHasher::build_merkle_tree (poseidon2.rs:297-316) */
int ba_poseidon2_gl_build_merkle_tree(const uint64_t *values, size_t values_len, uint64_t *results,
                                      size_t results_len, uint32_t log_rows_per_hash, uint32_t layers_count,
                                      ba_stream_t stream);

/* ---------------------------------------------------------------- launch tracing
 * No reference counterpart beyond its nvtx ranges (crates/shivini/Cargo.toml:31): after
 * ba_profile_begin(events, capacity) on a host thread, each kernel the library launches from
 * that thread is bracketed by events[2i] / events[2i+1] (caller-created) for its first
 * `capacity` launches; ba_profile_end reports how many were recorded and ba_profile_launch_name
 * the kernel name (as rocprofv3 prints it) of launch i. */
int ba_profile_begin(ba_event_t *events, uint32_t capacity);
int ba_profile_end(uint32_t *launches);
const char *ba_profile_launch_name(uint32_t index);

/* ---------------------------------------------------------------- helpers
 * Seeded uniform field elements on the device: element i of column k is the SplitMix64
 * finaliser of s_k + (i+1) * 0x9E3779B97F4A7C15 with s_k = seed + k * 0xD1B54A32D192ED03,
 * re-mixed while >= p.  No reference
 * counterpart (its benches draw unseeded host rng, benches/ntt.rs:70); used to build large
 * synthetic inputs in HBM without a host round trip. */
int ba_fill_uniform(uint64_t *dst, uint64_t stride, uint64_t rows, uint32_t cols, uint64_t seed,
                    ba_stream_t stream);

/* Element-wise field operation on the device (op: 0 add, 1 sub, 2 mul, 3 canonicalise,
 * 4 inverse), out[i] = op(a[i], b[i]).  Exposes the kernels' field arithmetic for parity
 * tests (goldilocks.cuh semantics, crates/boojum-cuda/native/goldilocks.cuh:90-317). */
int ba_field_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n,
                ba_stream_t stream);

/* Diagnostic, no reference counterpart: blocks x 256 threads each run `iters` x 8 butterflies
 * (one general product, one canonical add and subtract each, the NTT passes' arithmetic) on 8
 * register values seeded from in[thread], writing one word to out[thread] (in and out hold
 * blocks x 256 words).  Timing it gives the VALU issue rate of the passes' instruction mix
 * without memory, LDS or barriers. */
int ba_diag_butterfly_rate(const uint64_t *in, uint64_t *out, uint32_t blocks, uint32_t iters,
                           ba_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
