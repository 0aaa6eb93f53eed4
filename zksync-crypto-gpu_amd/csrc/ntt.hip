// ntt.hip -- batch NTT planner, launchers and the boojum_cuda::ntt entry points.
//
// Reference: crates/boojum-cuda/src/ntt.rs:57-369.  The reference plans 1-3 multi-stage
// kernels per log_n in 16..24 and falls back to log_n single-stage launches below 16
// (ntt.rs:201-241); here every size 1..30 runs in at most 1 + ceil((log_n - 12) / 8)
// passes of radix-16 register blocks (see ntt_kernels.h).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "api_internal.h"
#include "gl_host.h"
#include "ntt_inst.h"
#include "ntt_lat.h"

using gl::ntt::PassArgs;

bool ba_plan_from_bits(uint32_t log_n, bool n2b, const uint32_t *bits, int num_passes, NttPlan &plan) {
  constexpr uint32_t T = gl::ntt::TILE_LOG;
  if (num_passes < 1 || num_passes > 8) return false;
  if (log_n < 4) {  // whole transform per thread
    if (num_passes != 1 || bits[0] != log_n) return false;
  } else if (log_n <= T) {  // one contiguous pass
    if (num_passes != 1 || bits[0] != log_n) return false;
  } else {
    // contiguous pass over 9..12 low bits (whole 4096-element tiles), strided passes of 4..10
    if (num_passes < 2 || bits[0] < 9 || bits[0] > T) return false;
    uint32_t sum = bits[0];
    for (int i = 1; i < num_passes; i++) {
      if (bits[i] < 4 || bits[i] > 10) return false;
      sum += bits[i];
    }
    if (sum != log_n) return false;
  }
  plan = NttPlan{};
  plan.num_passes = num_passes;
  uint32_t lo[8];
  uint32_t at = 0;
  for (int i = 0; i < num_passes; i++) lo[i] = at, at += bits[i];
  for (int i = 0; i < num_passes; i++) {
    const int src = n2b ? num_passes - 1 - i : i;  // n2b: top bits first
    plan.lo_bit[i] = lo[src];
    plan.bits[i] = bits[src];
    plan.contig[i] = src == 0;
  }
  return true;
}

NttPlan ba_plan_ntt(uint32_t log_n, bool n2b) {
  // passes listed from the low bits up: a contiguous pass of 12 bits, then strided passes of
  // 8 bits, the last of 4..8 (bits borrowed from the pass below when fewer than 4 remain)
  uint32_t bits[8];
  int np = 0;
  if (log_n <= (uint32_t)gl::ntt::TILE_LOG) {
    bits[np++] = log_n;
  } else {
    uint32_t rest = log_n - gl::ntt::TILE_LOG;
    uint32_t sc = gl::ntt::TILE_LOG;
    if (rest < 4) {
      sc -= 4 - rest;
      rest = 4;
    }
    bits[np++] = sc;
    while (rest > 0) {
      uint32_t take = rest;
      if (rest > 8) take = (rest - 8 >= 4) ? 8 : rest - 4;
      bits[np++] = take;
      rest -= take;
    }
  }
  NttPlan plan{};
  (void)ba_plan_from_bits(log_n, n2b, bits, np, plan);
  return plan;
}

NttPlan ba_plan_for(const ba_context *ctx, uint32_t log_n, bool n2b) {
  const BaTuned &t = ctx->tuned[n2b][log_n];
  return t.valid ? t.plan : ba_plan_ntt(log_n, n2b);
}

namespace {

using namespace gl::ntt;

// scale: 0 none, 1 one factor for every element, 2 per-element coset factors (the strided
// kernels fuse 1 and 2 alike)
// waves: the 9..12-bit contiguous pass's variant (3 or 4 waves per SIMD)
// shared: the launch shares the GPU with other streams' work (strided passes: three-waves builds)
BaKernels pick_kernel(bool contig, uint32_t S, bool n2b, bool inverse, bool top, int scale,
                      bool whole, bool gather, uint32_t waves, bool shared) {
  if (contig) {
    if (S <= 8) return ba_pick_contig_lo(S, n2b, inverse, scale, whole);
    return waves == 3 ? ba_pick_contig_hi3(S, n2b, inverse, scale, whole)
                      : ba_pick_contig_hi(S, n2b, inverse, scale, whole);
  }
  return ba_pick_strided(S, n2b, inverse, top, scale != 0, gather, shared);
}

// kernel names as rocprofv3 prints them, for launch tracing; the table is built once, by the
// first caller (a function-local static initialiser, so concurrent first calls are safe)
struct KernelNames {
  // [strided, contig, strided shared build][S][n2b][inverse][top or whole][scale][gather or 4-waves]
  char n[3][13][2][2][2][3][2][120];
  KernelNames() {
    static const char *tf[2] = {"false", "true"};
    for (int s = 0; s < 13; s++)
      for (int d = 0; d < 2; d++)
        for (int i = 0; i < 2; i++)
          for (int tp = 0; tp < 2; tp++)
            for (int sc = 0; sc < 3; sc++)
              for (int ga = 0; ga < 2; ga++) {
                snprintf(n[1][s][d][i][tp][sc][ga], sizeof(n[1][s][d][i][tp][sc][ga]),
                         "void gl::ntt::ntt_pass_contig<%d, %s, %s, %d, %s, %d>(gl::ntt::PassArgs)", s, tf[d],
                         tf[i], sc, tf[tp], ga ? 4 : 3);
                snprintf(n[0][s][d][i][tp][sc][ga], sizeof(n[0][s][d][i][tp][sc][ga]),
                         "void gl::ntt::ntt_pass_strided<%d, %s, %s, %s, %s, %s>(gl::ntt::PassArgs)", s,
                         tf[d], tf[i], tf[tp], tf[sc != 0], tf[ga]);
                snprintf(n[2][s][d][i][tp][sc][ga], sizeof(n[2][s][d][i][tp][sc][ga]),
                         "void gl::ntt::ntt_pass_strided_shared<%d, %s, %s, %s, %s, %s>(gl::ntt::PassArgs)", s,
                         tf[d], tf[i], tf[tp], tf[sc != 0], tf[ga]);
              }
  }
};

// shared_build: a strided pass's three-waves build for shared launches
const char *kernel_name(bool contig, uint32_t S, bool n2b, bool inverse, bool top, int scale,
                        bool gather, uint32_t waves, bool shared_build) {
  static const KernelNames names;
  return names.n[contig ? 1 : (shared_build ? 2 : 0)][S][n2b][inverse][top][scale][contig ? waves == 4 : gather];
}

constexpr uint64_t kMaxGrid = 0x7FFFFFFFull;

}  // namespace

// Columns (tile repeats) per workgroup: enough to amortise the twiddle and scale-factor set-up,
// while keeping >= ~1024 workgroups (4 per CU) in flight.
uint32_t ba_pick_reps(uint64_t tiles, uint32_t cols) {
  uint32_t reps = 1;
  while (reps < 8 && reps < cols && tiles / (2ull * reps) >= 1024) reps *= 2;
  return reps;
}

// Latency-shaped plan (ntt_lat.h): the bits of each pass from the low bits up (a contiguous
// pass of <= 10 bits, then a strided pass of <= 8), or 0 when the size has none.
static int lat_bits(uint32_t log_n, uint32_t *bits) {
  if (log_n < 4 || log_n > 18) return 0;
  if (log_n <= 10) {
    bits[0] = log_n;
    return 1;
  }
  bits[1] = std::min<uint32_t>(8, log_n / 2);
  bits[0] = log_n - bits[1];
  return 2;
}

// BOOJUM_AMD_LAT_MAX_LOG_ELEMS overrides the batch-size bound (read once; 0 turns the latency
// passes off)
static uint32_t lat_max_log_elems() {
  static const uint32_t v = [] {
    const char *e = std::getenv("BOOJUM_AMD_LAT_MAX_LOG_ELEMS");
    return e && *e ? (uint32_t)std::strtoul(e, nullptr, 10) : BA_LAT_MAX_LOG_ELEMS;
  }();
  return v;
}

bool ba_latency_shape_exists(uint32_t log_n) {
  uint32_t bits[2];
  return lat_bits(log_n, bits) > 0;
}

bool ba_use_latency_passes(const ba_context *ctx, uint32_t log_n, bool n2b, uint32_t num_cols, BaLatPolicy lat) {
  if (lat == BA_LAT_NEVER || !ba_latency_shape_exists(log_n)) return false;
  // a tuned (or caller-set) tile choice for the size class, else the batch-size bound
  const BaTuned &t = ctx->tuned[n2b][log_n];
  if (t.valid && t.tile_log) return t.tile_log == LAT_TL;
  if (lat == BA_LAT_TUNED) return false;
  const uint32_t m = lat_max_log_elems();
  return m > 0 && ((uint64_t)num_cols << log_n) <= (1ull << std::min(m, 40u));
}

int ba_build_ntt(const ba_context *ctx, const uint64_t *in, uint64_t *out, uint32_t log_n,
                 uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs,
                 bool inverse, uint32_t log_ext, uint32_t coset_idx, bool in_bitrev_gather,
                 std::vector<BaPassLaunch> &launches, BaLatPolicy lat, bool shared) {
  if (num_cols == 0) return BA_SUCCESS;
  const bool n2b = !bitrev_inputs;
  PassArgs base{};
  base.log_n = log_n;
  base.scale_coset = log_ext > 0;
  base.coset_mul = log_ext > 0 ? coset_idx << (gl::OMEGA_LOG_ORDER - log_n - log_ext) : 0;
  base.tab = ctx->tab;
  const bool pre = !inverse && log_ext > 0;
  // scale table T[m] = Delta^m (times n^-1 for the inverse), Delta = s^(+-2^(L-4))
  {
    const uint64_t s = gl_host::coset_shift(log_n, log_ext, coset_idx);
    const uint64_t n_inv = gl_host::inv(1ull << log_n);
    const uint64_t sd = log_n >= 4 ? gl_host::pow(s, 1ull << (log_n - 4)) : 1;
    const uint64_t delta = inverse ? gl_host::inv(sd) : sd;
    uint64_t acc = inverse ? n_inv : 1;
    for (int m = 0; m < VALS; m++) {
      base.scale_tab[m] = acc;
      acc = gl_host::mul(acc, delta);
    }
  }

  if (log_n < 4) {
    BaPassLaunch l;
    l.args = base;
    l.args.in = in;
    l.args.out = out;
    l.args.stride_in = stride_in;
    l.args.stride_out = stride_out;
    l.args.num_cols = num_cols;
    l.args.scale = pre ? SCALE_PRE : (inverse ? SCALE_POST : SCALE_NONE);
    l.args.in_bitrev = in_bitrev_gather;
    l.kernels = ba_pick_tiny(log_n, n2b, inverse);
    l.grid = (num_cols + THREADS - 1) / THREADS;
    l.name = "ntt_tiny";
    l.pass = 0;
    l.col0 = 0;
    launches.push_back(l);
    return BA_SUCCESS;
  }

  if (!in_bitrev_gather && ba_use_latency_passes(ctx, log_n, n2b, num_cols, lat)) {
    uint32_t bits[2];
    const int np = lat_bits(log_n, bits);
    for (int p = 0; p < np; p++) {
      const int src = n2b ? np - 1 - p : p;  // n2b: top bits first
      const uint32_t S = bits[src], lo = src == 0 ? 0 : bits[0];
      const bool strided = src > 0, top = lo + S == log_n;
      BaPassLaunch l;
      PassArgs &a = l.args;
      a = base;
      a.in = p == 0 ? in : out;
      a.stride_in = p == 0 ? stride_in : stride_out;
      a.out = out;
      a.stride_out = stride_out;
      a.num_cols = num_cols;
      a.lo_bit = lo;
      a.reps = 1;
      a.in_canonical = p > 0;
      a.scale = (pre && p == 0) ? SCALE_PRE : ((inverse && p == np - 1) ? SCALE_POST : SCALE_NONE);
      const bool scale = a.scale != SCALE_NONE;
      const uint64_t chunks = (uint64_t)num_cols << (log_n - S);
      l.grid = strided ? ((uint64_t)num_cols << (log_n - LAT_TL))
                       : (chunks + (1u << (LAT_TL - S)) - 1) >> (LAT_TL - S);
      a.tile_count = (uint32_t)l.grid;
      l.kernels = ba_pick_lat(S, strided, n2b, inverse, top, scale);
      l.name = ba_lat_kernel_name(S, strided, n2b, inverse, top, scale);
      l.pass = p;
      l.col0 = 0;
      if (!l.kernels.single) return BA_ERROR_INVALID_VALUE;
      launches.push_back(l);
    }
    return BA_SUCCESS;
  }

  // The size class's tuned plan, else the built-in one.  An LDE's coset NTTs (the gathered n2b
  // forward) have their own entry, timed on that transform: a plan tuned on plain NTTs does not
  // carry over (the gathered, scaled first pass has other costs; it has no 9- or 10-bit variant).
  const int kind = in_bitrev_gather ? BA_TUNING_LDE_COSETS : (n2b ? 1 : 0);
  const BaTuned &tuning = ctx->tuned[kind][log_n].valid ? ctx->tuned[kind][log_n] : BaTuned{};
  const NttPlan plan = tuning.valid ? tuning.plan : ba_plan_ntt(log_n, n2b);
  for (int p = 0; p < plan.num_passes; p++) {
    const uint32_t S = plan.bits[p];
    const bool contig = plan.contig[p];
    // tiles: L >= 12 -> 2^(L-12) per column, repeats stride over columns;
    //        L < 12  -> 2^(12-L) columns per tile (contiguous pass only), repeats adjacent
    const bool big = log_n >= (uint32_t)TILE_LOG;
    const uint32_t log_stride = big ? log_n - TILE_LOG : 0;
    const uint32_t cols_per_tile = big ? 1 : (1u << (TILE_LOG - log_n));
    // launch column groups small enough for 32-bit tile counts and grids
    const uint64_t max_cols = big ? std::max<uint64_t>(1, (kMaxGrid >> log_stride)) : kMaxGrid;
    for (uint64_t c0 = 0; c0 < num_cols; c0 += max_cols) {
      const uint32_t nc = (uint32_t)std::min<uint64_t>(max_cols, num_cols - c0);
      const uint64_t tiles = big ? ((uint64_t)nc << log_stride) : (nc + cols_per_tile - 1) / cols_per_tile;
      BaPassLaunch l;
      PassArgs &a = l.args;
      a = base;
      a.in = p == 0 ? in + c0 * stride_in : out + c0 * stride_out;
      a.stride_in = p == 0 ? stride_in : stride_out;
      a.out = out + c0 * stride_out;
      a.stride_out = stride_out;
      a.num_cols = nc;
      a.lo_bit = plan.lo_bit[p];
      a.tile_count = (uint32_t)tiles;
      a.log_tile_stride = log_stride;
      const uint32_t tuned_reps = tuning.valid ? tuning.reps[p] : 0;
      a.reps = tuned_reps ? std::min<uint32_t>(tuned_reps, big ? nc : (uint32_t)tiles)
                     : ba_pick_reps(tiles, big ? nc : (uint32_t)tiles);
      if (a.reps == 0) a.reps = 1;
      a.scale = (pre && p == 0) ? SCALE_PRE
                                : ((inverse && p == plan.num_passes - 1) ? SCALE_POST : SCALE_NONE);
      a.in_bitrev = in_bitrev_gather && p == 0;
      // later passes read this transform's own output, and the gathered LDE coset pass reads the
      // monomials the inverse NTT wrote: values < p
      a.in_canonical = p > 0 || in_bitrev_gather;
      const uint64_t groups = big ? (nc + a.reps - 1) / a.reps : (tiles + a.reps - 1) / a.reps;
      l.grid = groups << log_stride;
      const bool top = plan.lo_bit[p] + S == log_n;
      const int scale = a.scale == SCALE_NONE ? 0 : (a.scale_coset ? 2 : 1);
      const bool gather = !contig && a.in_bitrev;
      // the 9..12-bit contiguous pass: as tuned, else 3 waves per SIMD (with the next-tile
      // prefetch) below 2048 workgroups -- few tiles, where a workgroup's own instruction stream
      // sets the time -- and 4 waves above
      uint32_t waves = 3;
      if (contig && S >= 9)
        waves = tuning.valid && tuning.contig_waves ? tuning.contig_waves : (l.grid < 2048 ? 3 : 4);
      l.kernels = pick_kernel(contig, S, n2b, inverse, top, scale, big, gather, waves, shared);
      const bool shared_build =
          shared && !contig && l.kernels.single != pick_kernel(contig, S, n2b, inverse, top, scale, big, gather, waves, false).single;
      l.name = kernel_name(contig, S, n2b, inverse, contig ? big : top, scale, gather, waves, shared_build);
      l.pass = p;
      l.col0 = c0;
      if (!l.kernels.single) return BA_ERROR_INVALID_VALUE;
      launches.push_back(l);
    }
  }
  return BA_SUCCESS;
}

int ba_run_launches(const std::vector<BaPassLaunch> &launches, hipStream_t stream) {
  for (const BaPassLaunch &l : launches) {
    ba_trace_before(stream, l.name);
    hipLaunchKernelGGL(l.kernels.single, dim3((uint32_t)l.grid), dim3(THREADS), 0, stream, l.args);
    ba_trace_after(stream);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return (int)err;
  }
  return BA_SUCCESS;
}

int ba_launch_ntt(const ba_context *ctx, const uint64_t *in, uint64_t *out, uint32_t log_n,
                  uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs,
                  bool inverse, uint32_t log_ext, uint32_t coset_idx, hipStream_t stream,
                  bool in_bitrev_gather) {
  std::vector<BaPassLaunch> launches;
  const int rc = ba_build_ntt(ctx, in, out, log_n, num_cols, stride_in, stride_out, bitrev_inputs,
                              inverse, log_ext, coset_idx, in_bitrev_gather, launches, BA_LAT_AUTO);
  if (rc) return rc;
  return ba_run_launches(launches, stream);
}

int ba_validate_ntt(uint32_t log_n, uint64_t stride_in, uint64_t stride_out, uint32_t log_ext,
                    uint32_t coset_idx) {
  // ntt.rs:192-200; in addition a column holds at most 2^28 elements (the reference's plans
  // stop at 2^24, ntt.rs:74-169, and index past them for larger sizes)
  if (log_n < 1 || log_n > BA_MAX_LOG_N || log_n + log_ext > gl::OMEGA_LOG_ORDER)
    return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  if (n > stride_in || n > stride_out) return BA_ERROR_INVALID_VALUE;
  if (coset_idx > 0 && (log_ext == 0 || coset_idx >= (1u << log_ext))) return BA_ERROR_INVALID_VALUE;
  return BA_SUCCESS;
}

extern "C" int ba_batch_ntt_ctx(ba_context_t ctx, const uint64_t *inputs, uint64_t *outputs,
                                uint32_t log_n, uint32_t num_ntts, uint64_t stride_in,
                                uint64_t stride_out, int bitrev_inputs, int inverse,
                                uint32_t log_extension_degree, uint32_t coset_index,
                                ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    int rc = ba_validate_ntt(log_n, stride_in, stride_out, log_extension_degree, coset_index);
    if (rc) return rc;
    if (num_ntts && (!inputs || !outputs)) return BA_ERROR_INVALID_VALUE;
    return ba_launch_ntt(ctx, inputs, outputs, log_n, num_ntts, stride_in, stride_out,
                         bitrev_inputs != 0, inverse != 0, log_extension_degree, coset_index,
                         (hipStream_t)stream);
  });
}

extern "C" int ba_batch_ntt_internal(const uint64_t *inputs, uint64_t *outputs, uint32_t log_n,
                                     uint32_t num_ntts, uint32_t stride_between_input_arrays,
                                     uint32_t stride_between_output_arrays, int bitrev_inputs,
                                     int inverse, uint32_t log_extension_degree,
                                     uint32_t coset_index, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    return ba_batch_ntt_ctx(ba_context_for_current_device(), inputs, outputs, log_n, num_ntts,
                            stride_between_input_arrays, stride_between_output_arrays,
                            bitrev_inputs, inverse, log_extension_degree, coset_index, stream);
  });
}

extern "C" int ba_batch_ntt_out_of_place(const uint64_t *inputs_matrix, size_t inputs_len,
                                         uint64_t *outputs_matrix, size_t outputs_len,
                                         uint32_t log_n, uint32_t num_ntts, uint32_t inputs_offset,
                                         uint32_t outputs_offset,
                                         uint32_t stride_between_input_arrays,
                                         uint32_t stride_between_output_arrays, int bitrev_inputs,
                                         int inverse, uint32_t log_extension_degree,
                                         uint32_t coset_index, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    // ntt.rs:309-320, with 64-bit arithmetic in place of the reference's u32 products
    if (log_n < 1 || log_n > BA_MAX_LOG_N) return BA_ERROR_INVALID_VALUE;
    const uint64_t n = 1ull << log_n;
    if ((uint64_t)inputs_offset + n > stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
    if ((uint64_t)outputs_offset + n > stride_between_output_arrays) return BA_ERROR_INVALID_VALUE;
    if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
    if ((uint64_t)outputs_len < (uint64_t)num_ntts * stride_between_output_arrays) return BA_ERROR_INVALID_VALUE;
    // the reference's slices are distinct borrows (ntt.rs:317,320): no overlap at all
    if (inputs_len && outputs_len && inputs_matrix < outputs_matrix + outputs_len &&
        outputs_matrix < inputs_matrix + inputs_len)
      return BA_ERROR_INVALID_VALUE;
    const uint64_t *in = inputs_matrix + inputs_offset;
    uint64_t *out = outputs_matrix + outputs_offset;
    if ((const void *)in == (const void *)out) return BA_ERROR_INVALID_VALUE;
    return ba_batch_ntt_internal(in, out, log_n, num_ntts, stride_between_input_arrays,
                                 stride_between_output_arrays, bitrev_inputs, inverse,
                                 log_extension_degree, coset_index, stream);
  });
}

extern "C" int ba_batch_ntt_in_place(uint64_t *inputs_matrix, size_t inputs_len, uint32_t log_n,
                                     uint32_t num_ntts, uint32_t inputs_offset,
                                     uint32_t stride_between_input_arrays, int bitrev_inputs,
                                     int inverse, uint32_t log_extension_degree,
                                     uint32_t coset_index, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    // ntt.rs:350-352
    if (log_n < 1 || log_n > BA_MAX_LOG_N) return BA_ERROR_INVALID_VALUE;
    const uint64_t n = 1ull << log_n;
    if ((uint64_t)inputs_offset + n > stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
    if ((uint64_t)inputs_len < (uint64_t)num_ntts * stride_between_input_arrays) return BA_ERROR_INVALID_VALUE;
    uint64_t *io = inputs_matrix + inputs_offset;
    return ba_batch_ntt_internal(io, io, log_n, num_ntts, stride_between_input_arrays,
                                 stride_between_input_arrays, bitrev_inputs, inverse,
                                 log_extension_degree, coset_index, stream);
  });
}
