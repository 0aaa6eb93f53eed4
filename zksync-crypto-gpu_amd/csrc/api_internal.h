// api_internal.h -- definitions shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <new>

#include "../../include/boojum_amd.h"
#include "gl_tables.h"

// Runs the body of an extern "C" entry point so that no C++ exception crosses the C-ABI: host
// allocation failures (planning vectors, server state, threads) become BA_ERROR_OUT_OF_MEMORY,
// anything else BA_ERROR_INVALID_VALUE.
template <typename F>
int ba_guarded(F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return BA_ERROR_OUT_OF_MEMORY;
  } catch (...) {
    return BA_ERROR_INVALID_VALUE;
  }
}

// Launch-time plan of one NTT: the index bits each pass transforms, in execution order.
struct NttPlan {
  int num_passes;
  uint32_t lo_bit[8];
  uint32_t bits[8];
  bool contig[8];
};

// An autotuned (or caller-set) choice for one size class: the pass split and, per pass in
// execution order, the columns (tiles) each workgroup repeats over (0 = built-in choice).
struct BaTuned {
  bool valid;
  NttPlan plan;
  uint32_t reps[8];
  uint32_t contig_waves;  // waves per SIMD of the 9..12-bit contiguous pass: 3, 4, or 0 = built-in
  uint32_t tile_log;      // plain batch NTTs: 10 latency-shaped tiles, 12 4096-element tiles, 0 by batch size
};

struct ba_context {
  int device;
  uint32_t w_coarse_log;
  uint32_t g_coarse_log;
  void *mem;
  size_t bytes;
  gl::Tables tab;
  // [kind][log_n]: kind 0 plain b2n, 1 plain n2b (and the LDE's inverse NTT), 2 the LDE's
  // gathered coset NTTs (BA_TUNING_LDE_COSETS)
  BaTuned tuned[3][BA_MAX_LOG_N + 1] = {};
  // batch plans and servers hold pointers into the tables: the context outlives them
  std::atomic<int> users{0};
  // Poseidon2 round constants (30 x 12, device memory), installed by the caller
  uint64_t *p2_round_constants = nullptr;
  // ba_lde_batch runs the coset NTTs after the first on these streams (created with the context),
  // forked from and joined back into the caller's stream; the mutex keeps two callers' fork and
  // join events from interleaving
  static constexpr int kLdeAux = 7;
  hipStream_t lde_aux[kLdeAux] = {};
  hipEvent_t lde_fork = nullptr, lde_join[kLdeAux] = {};
  std::mutex lde_mutex;
};

// The most recently created live context of `device`, or null.
ba_context *ba_current_context(int device);

// Context of the current device for the context-free entry points.
inline ba_context *ba_context_for_current_device() {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return nullptr;
  return ba_current_context(device);
}

// Pass decomposition for a 2^log_n transform (n2b order: strided passes from the top bits
// down, then the contiguous pass; b2n runs the same passes in reverse).
NttPlan ba_plan_ntt(uint32_t log_n, bool n2b);
// The plan whose passes cover bits[0] (contiguous pass, lowest bits), bits[1], ... upwards;
// false when no kernel shape covers that split.
bool ba_plan_from_bits(uint32_t log_n, bool n2b, const uint32_t *bits, int num_passes, NttPlan &plan);
// The plan a transform of this size class runs with on `ctx` (tuned or built-in).
NttPlan ba_plan_for(const ba_context *ctx, uint32_t log_n, bool n2b);

// One pass launch of a batch NTT: the kernel pair of its shape, its arguments and grid.
#include <utility>
#include <vector>
#include "ntt_inst.h"
struct BaPassLaunch {
  BaKernels kernels;
  gl::ntt::PassArgs args;
  uint64_t grid;
  const char *name;
  int pass;       // index of the pass inside its transform
  uint64_t col0;  // first column of the transform this launch covers
};

// Which transforms may run the latency-shaped passes (1024-element tiles):
//   BA_LAT_NEVER  LDE transforms (their coset NTTs gather, and their inverse NTT keeps the
//                 cosets' pass count);
//   BA_LAT_TUNED  the size class's tuned tile_log is 10 (batch plans: a merged small transform
//                 runs beside other groups of its level, where the throughput passes measured
//                 as fast and the extra grouped launches cost more);
//   BA_LAT_AUTO   that, or else the batch holds at most 2^BA_LAT_MAX_LOG_ELEMS elements and the
//                 tuned tile_log is not 12 (single batch NTT calls).
enum BaLatPolicy { BA_LAT_NEVER = 0, BA_LAT_TUNED = 1, BA_LAT_AUTO = 2 };
// Appends the pass launches of one batch NTT (arguments validated) to `launches`.  shared: the
// launches will run beside other streams' work (serve batches), so strided passes whose default
// build runs four waves per SIMD, filling a CU's LDS, take their three-waves build.
int ba_build_ntt(const ba_context *ctx, const uint64_t *in, uint64_t *out, uint32_t log_n,
                 uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs,
                 bool inverse, uint32_t log_ext, uint32_t coset_idx, bool in_bitrev_gather,
                 std::vector<BaPassLaunch> &launches, BaLatPolicy lat = BA_LAT_NEVER, bool shared = false);
// Per-call time over log_n 10..18 and batches of 2^14..2^22 elements (MI355X): up to 2^20
// elements the latency passes take 1.0-3.1x less time per call, above it the throughput passes win.
constexpr uint32_t BA_LAT_MAX_LOG_ELEMS = 20;
// log_n has a latency-shaped plan (4..18); a batch NTT of num_cols columns of 2^log_n runs it.
bool ba_latency_shape_exists(uint32_t log_n);
bool ba_use_latency_passes(const ba_context *ctx, uint32_t log_n, bool n2b, uint32_t num_cols, BaLatPolicy lat);
// Launches them one after another on `stream` (single-batch kernels).
int ba_run_launches(const std::vector<BaPassLaunch> &launches, hipStream_t stream);
uint32_t ba_pick_reps(uint64_t tiles, uint32_t cols);

// A pass launch tagged with its dependency level inside a batch plan: every launch at level l
// only reads what launches at levels < l wrote.
struct BaLeveledLaunch {
  uint32_t level;
  BaPassLaunch launch;
};
// A batch plan as data: grouped launches (one per dependency level and pass shape) and the
// tables they read, laid out for one device allocation of `bytes` bytes.
struct BaGroup {
  BaGroupKernel kernel;
  const char *name;
  uint32_t level;
  uint32_t njobs;
  uint64_t workgroups;
  size_t args_off;    // index into the PassArgs table
  size_t prefix_off;  // index into the prefix table
};
struct BaPlanImage {
  std::vector<BaGroup> groups;  // ordered by level, largest first within a level
  uint32_t levels = 0;
  uint64_t workgroups = 0;
  std::vector<gl::ntt::PassArgs> args;
  std::vector<std::pair<size_t, size_t>> lists;  // (in, out) offsets into `table` per PassArgs
  std::vector<const void *> table;               // column pointers
  std::vector<uint32_t> prefix;                  // first workgroup of each PassArgs in its group
  size_t table_at = 0, prefix_at = 0, bytes = 0;
};
// shared: the plan's launches run beside other streams' work (serve batches; see ba_build_ntt)
int ba_plan_image_build(const ba_context *ctx, const ba_job *jobs, uint32_t num_jobs, BaPlanImage &img,
                        bool shared = false);
// Serialises the image into `host` (img.bytes bytes) for a copy to device address `device`.
void ba_plan_image_write(const BaPlanImage &img, void *host, void *device);
// Enqueues one group of an image uploaded at `device`.
hipError_t ba_plan_launch_group(const BaGroup &g, const void *device, const BaPlanImage &img, hipStream_t s);

int ba_validate_lde(uint64_t *evals, uint64_t in_stride, uint64_t *monomials, uint64_t *out,
                    uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                    uint32_t num_cols, uint32_t log_blowup);
int ba_build_lde(const ba_context *ctx, uint64_t *evals, uint64_t in_stride, uint64_t *monomials,
                 uint64_t *out, uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                 uint32_t num_cols, uint32_t log_blowup, std::vector<BaLeveledLaunch> &launches,
                 uint32_t first_level, size_t *inverse_launches = nullptr);
int ba_validate_ntt(uint32_t log_n, uint64_t stride_in, uint64_t stride_out, uint32_t log_ext,
                    uint32_t coset_idx);

// Enqueues one batch NTT (arguments already validated).  in_bitrev_gather (natural -> bitrev
// forward only): the input holds each column in bit-reversed order and the first pass gathers
// it, so a bit-reversed monomial vector feeds a coset NTT without a separate permutation.
int ba_launch_ntt(const ba_context *ctx, const uint64_t *in, uint64_t *out, uint32_t log_n,
                  uint32_t num_cols, uint64_t stride_in, uint64_t stride_out, bool bitrev_inputs,
                  bool inverse, uint32_t log_ext, uint32_t coset_idx, hipStream_t stream,
                  bool in_bitrev_gather = false);

// Launch tracing: when a host thread has called ba_profile_begin, every kernel the library
// launches from that thread is bracketed by two of the caller's events.
void ba_trace_before(hipStream_t stream, const char *kernel);
void ba_trace_after(hipStream_t stream);
// launch tracing is on for this host thread (per-launch timings need launches that do not overlap)
bool ba_trace_active();
