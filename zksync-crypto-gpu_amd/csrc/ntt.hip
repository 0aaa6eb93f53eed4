// ntt.hip -- batch NTT planner, launchers and the boojum_cuda::ntt entry points.
//
// Reference: crates/boojum-cuda/src/ntt.rs:57-369.  The reference plans 1-3 multi-stage
// kernels per log_n in 16..24 and falls back to log_n single-stage launches below 16
// (ntt.rs:201-241); here every size 1..30 runs in at most 1 + ceil((log_n - 12) / 8)
// passes of radix-16 register blocks (see ntt_kernels.h).
#include <algorithm>

#include "api_internal.h"
#include "ntt_kernels.h"

using gl::ntt::PassArgs;

NttPlan ba_plan_ntt(uint32_t log_n, bool n2b) {
  NttPlan plan{};
  // passes listed from the low bits up
  uint32_t lo[8], bits[8];
  bool contig[8];
  int np = 0;
  if (log_n <= (uint32_t)gl::ntt::TILE_LOG) {
    lo[np] = 0, bits[np] = log_n, contig[np] = true, np++;
  } else {
    uint32_t rest = log_n - gl::ntt::TILE_LOG;
    uint32_t sc = gl::ntt::TILE_LOG;
    if (rest < 4) {  // the strided pass needs >= 4 bits: borrow them from the contiguous pass
      sc -= 4 - rest;
      rest = 4;
    }
    lo[np] = 0, bits[np] = sc, contig[np] = true, np++;
    uint32_t at = sc;
    while (rest > 0) {
      uint32_t take = rest;
      if (rest > 8) take = (rest - 8 >= 4) ? 8 : rest - 4;
      lo[np] = at, bits[np] = take, contig[np] = false, np++;
      at += take;
      rest -= take;
    }
  }
  plan.num_passes = np;
  for (int i = 0; i < np; i++) {
    const int src = n2b ? np - 1 - i : i;  // n2b: top bits first
    plan.lo_bit[i] = lo[src];
    plan.bits[i] = bits[src];
    plan.contig[i] = contig[src];
  }
  return plan;
}

namespace {

typedef void (*PassKernel)(PassArgs);

PassKernel pick_kernel(bool contig, bool n2b, bool inverse) {
  using namespace gl::ntt;
  if (contig) {
    if (n2b) return inverse ? ntt_pass_contig<true, true> : ntt_pass_contig<true, false>;
    return inverse ? ntt_pass_contig<false, true> : ntt_pass_contig<false, false>;
  }
  if (n2b) return inverse ? ntt_pass_strided<true, true> : ntt_pass_strided<true, false>;
  return inverse ? ntt_pass_strided<false, true> : ntt_pass_strided<false, false>;
}

PassKernel pick_tiny(uint32_t log_n, bool n2b, bool inverse) {
  using namespace gl::ntt;
#define BA_TINY(K)                                                        \
  if (log_n == K) {                                                       \
    if (n2b) return inverse ? ntt_tiny<K, true, true> : ntt_tiny<K, true, false>; \
    return inverse ? ntt_tiny<K, false, true> : ntt_tiny<K, false, false>;        \
  }
  BA_TINY(1)
  BA_TINY(2)
  BA_TINY(3)
#undef BA_TINY
  return nullptr;
}

constexpr uint64_t kMaxGrid = 0x7FFFFFFFull;

// names as rocprofv3 reports the kernels
const char *kernel_name(bool contig, bool n2b, bool inverse) {
  static const char *names[8] = {
      "void gl::ntt::ntt_pass_strided<false, false>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_strided<false, true>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_strided<true, false>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_strided<true, true>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_contig<false, false>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_contig<false, true>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_contig<true, false>(gl::ntt::PassArgs)",
      "void gl::ntt::ntt_pass_contig<true, true>(gl::ntt::PassArgs)",
  };
  return names[(contig ? 4 : 0) + (n2b ? 2 : 0) + (inverse ? 1 : 0)];
}

}  // namespace

int ba_launch_ntt(const ba_context *ctx, const uint64_t *in, uint64_t *out, uint32_t log_n,
                  uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs,
                  bool inverse, uint32_t log_ext, uint32_t coset_idx, hipStream_t stream,
                  bool in_bitrev_gather) {
  if (num_cols == 0) return BA_SUCCESS;
  const bool n2b = !bitrev_inputs;
  PassArgs base{};
  base.log_n = log_n;
  base.coset = log_ext > 0;
  base.coset_mul = log_ext > 0 ? coset_idx << (gl::OMEGA_LOG_ORDER - log_n - log_ext) : 0;
  base.tab = ctx->tab;
  const bool pre = !inverse && log_ext > 0;

  if (log_n < 4) {
    PassArgs a = base;
    a.in = in;
    a.out = out;
    a.stride_in = stride_in;
    a.stride_out = stride_out;
    a.num_cols = num_cols;
    a.bits = log_n;
    a.pre_scale = pre;
    a.post_scale = inverse;
    a.in_bitrev = in_bitrev_gather;
    const uint32_t grid = (num_cols + gl::ntt::THREADS - 1) / gl::ntt::THREADS;
    ba_trace_before(stream, "ntt_tiny");
    hipLaunchKernelGGL(pick_tiny(log_n, n2b, inverse), dim3(grid), dim3(gl::ntt::THREADS), 0,
                       stream, a);
    ba_trace_after(stream);
    return (int)hipGetLastError();
  }

  const NttPlan plan = ba_plan_ntt(log_n, n2b);
  for (int p = 0; p < plan.num_passes; p++) {
    const uint32_t S = plan.bits[p];
    const bool contig = plan.contig[p];
    // tiles per column (strided and large contiguous passes) or columns per tile (small)
    const uint64_t per_col_tiles = log_n >= (uint32_t)gl::ntt::TILE_LOG
                                       ? (1ull << (log_n - gl::ntt::TILE_LOG))
                                       : 0;
    const uint32_t cols_per_tile = per_col_tiles ? 1 : (1u << (gl::ntt::TILE_LOG - log_n));
    const uint64_t max_cols = per_col_tiles ? std::max<uint64_t>(1, kMaxGrid / per_col_tiles)
                                            : kMaxGrid * cols_per_tile;
    for (uint64_t c0 = 0; c0 < num_cols; c0 += max_cols) {
      const uint32_t nc = (uint32_t)std::min<uint64_t>(max_cols, num_cols - c0);
      PassArgs a = base;
      a.in = p == 0 ? in + c0 * stride_in : out + c0 * stride_out;
      a.stride_in = p == 0 ? stride_in : stride_out;
      a.out = out + c0 * stride_out;
      a.stride_out = stride_out;
      a.num_cols = nc;
      a.lo_bit = plan.lo_bit[p];
      a.bits = S;
      a.pre_scale = pre && p == 0;
      a.post_scale = inverse && p == plan.num_passes - 1;
      a.in_bitrev = in_bitrev_gather && p == 0;
      const uint64_t grid = per_col_tiles ? per_col_tiles * nc : (nc + cols_per_tile - 1) / cols_per_tile;
      ba_trace_before(stream, kernel_name(contig, n2b, inverse));
      hipLaunchKernelGGL(pick_kernel(contig, n2b, inverse), dim3((uint32_t)grid),
                         dim3(gl::ntt::THREADS), 0, stream, a);
      ba_trace_after(stream);
      const hipError_t err = hipGetLastError();
      if (err != hipSuccess) return (int)err;
    }
  }
  return BA_SUCCESS;
}

namespace {

int validate(uint32_t log_n, uint64_t stride_in, uint64_t stride_out, uint32_t log_ext,
             uint32_t coset_idx) {
  // ntt.rs:192-200
  if (log_n < 1 || log_n + log_ext > gl::OMEGA_LOG_ORDER) return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if (n > stride_in || n > stride_out) return BA_ERROR_INVALID_VALUE;
  if (coset_idx > 0 && (log_ext == 0 || coset_idx >= (1u << log_ext))) return BA_ERROR_INVALID_VALUE;
  return BA_SUCCESS;
}

}  // namespace

extern "C" int ba_batch_ntt_ctx(ba_context_t ctx, const uint64_t *inputs, uint64_t *outputs,
                                uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                                uint64_t stride_out, int bitrev_inputs, int inverse,
                                uint32_t log_extension_degree, uint32_t coset_index,
                                ba_stream_t stream) {
  if (!ctx) return BA_ERROR_NOT_INITIALIZED;
  int rc = validate(log_n, stride_in, stride_out, log_extension_degree, coset_index);
  if (rc) return rc;
  if (num_ntts && (!inputs || !outputs)) return BA_ERROR_INVALID_VALUE;
  return ba_launch_ntt(ctx, inputs, outputs, log_n, num_ntts, stride_in, stride_out,
                       bitrev_inputs != 0, inverse != 0, log_extension_degree, coset_index,
                       (hipStream_t)stream);
}

extern "C" int ba_batch_ntt_internal(const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                                     uint32_t num_ntts, uint32_t stride_between_input_arrays,
                                     uint32_t stride_between_output_arrays, int bitrev_inputs,
                                     int inverse, uint32_t log_extension_degree,
                                     uint32_t coset_index, ba_stream_t stream) {
  return ba_batch_ntt_ctx(ba_context_for_current_device(), inputs, outputs, log_n, num_ntts,
                          stride_between_input_arrays, stride_between_output_arrays,
                          bitrev_inputs, inverse, log_extension_degree, coset_index, stream);
}

extern "C" int ba_batch_ntt_out_of_place(const uint64_t *inputs_matrix, size_t inputs_len,
                                         uint64_t *outputs_matrix, size_t outputs_len,
                                         uint32_t log_n, uint32_t num_ntts, uint32_t inputs_offset,
                                         uint32_t outputs_offset,
                                         uint32_t stride_between_input_arrays,
                                         uint32_t stride_between_output_arrays, int bitrev_inputs,
                                         int inverse, uint32_t log_extension_degree,
                                         uint32_t coset_index, ba_stream_t stream) {
  // ntt.rs:309-320, with 64-bit arithmetic in place of the reference's u32 products
  if (log_n < 1 || log_n > gl::OMEGA_LOG_ORDER) return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if ((uint64_t)inputs_offset + n > stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
  if ((uint64_t)outputs_offset + n > stride_between_output_arrays) return BA_ERROR_INVALID_VALUE;
  if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
  if ((uint64_t)outputs_len < (uint64_t)num_ntts * stride_between_output_arrays) return BA_ERROR_INVALID_VALUE;
  if ((const void *)inputs_matrix == (const void *)outputs_matrix) return BA_ERROR_INVALID_VALUE;
  const uint64_t *in = inputs_matrix + inputs_offset;
  uint64_t *out = outputs_matrix + outputs_offset;
  if ((const void *)in == (const void *)out) return BA_ERROR_INVALID_VALUE;
  return ba_batch_ntt_internal(in, out, log_n, num_ntts, stride_between_input_arrays,
                               stride_between_output_arrays, bitrev_inputs, inverse,
                               log_extension_degree, coset_index, stream);
}

extern "C" int ba_batch_ntt_in_place(uint64_t *inputs_matrix, size_t inputs_len, uint32_t log_n,
                                     uint32_t num_ntts, uint32_t inputs_offset,
                                     uint32_t stride_between_input_arrays, int bitrev_inputs,
                                     int inverse, uint32_t log_extension_degree,
                                     uint32_t coset_index, ba_stream_t stream) {
  // ntt.rs:350-352
  if (log_n < 1 || log_n > gl::OMEGA_LOG_ORDER) return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if ((uint64_t)inputs_offset + n > stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
  if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
  uint64_t *io = inputs_matrix + inputs_offset;
  return ba_batch_ntt_internal(io, io, log_n, num_ntts, stride_between_input_arrays,
                               stride_between_input_arrays, bitrev_inputs, inverse,
                               log_extension_degree, coset_index, stream);
}
