// tune.hip -- the autotuner: per size class (log_n, input order) it times every pass split, then
// the columns-per-workgroup of each pass, then the waves per SIMD the contiguous pass runs at,
// then the tile shape (4096-element tiles of 16 elements per thread against the latency-shaped
// 1024-element tiles of 4), and keeps the fastest on the context.
//
// No reference counterpart: boojum-cuda's kernel shapes per log_n are fixed tables
// (crates/boojum-cuda/src/ntt.rs:57-176).  The candidate splits are those the pass kernels
// cover: one contiguous pass for log_n <= 12; otherwise a contiguous pass over 9..12 low bits
// (whole 4096-element tiles) below strided passes of 4..10 bits each, from the fewest strided
// passes that cover the remaining bits to one more than that (an extra pass adds HBM traffic
// but can pay for itself when the passes are bound by integer ALU work).
#include <algorithm>
#include <vector>

#include "api_internal.h"
#include "ntt_lat.h"

namespace {

// every composition of `rest` into `parts` pieces of 4..10 bits
void compositions(uint32_t rest, int parts, std::vector<uint32_t> &cur,
                  std::vector<std::vector<uint32_t>> &out) {
  if (parts == 0) {
    if (rest == 0) out.push_back(cur);
    return;
  }
  for (uint32_t b = 4; b <= 10 && b <= rest; b++) {
    cur.push_back(b);
    compositions(rest - b, parts - 1, cur, out);
    cur.pop_back();
  }
}

// candidate splits, bits listed from the low (contiguous) pass up
std::vector<std::vector<uint32_t>> candidate_splits(uint32_t log_n) {
  std::vector<std::vector<uint32_t>> out;
  if (log_n <= (uint32_t)gl::ntt::TILE_LOG) {
    out.push_back({log_n});
    return out;
  }
  for (uint32_t sc = 9; sc <= (uint32_t)gl::ntt::TILE_LOG; sc++) {
    if (log_n < sc + 4) continue;
    const uint32_t rest = log_n - sc;
    const int min_parts = (int)((rest + 9) / 10);
    for (int parts = min_parts; parts <= min_parts + 1 && parts <= 7; parts++) {
      std::vector<std::vector<uint32_t>> comps;
      std::vector<uint32_t> cur;
      compositions(rest, parts, cur, comps);
      for (auto &c : comps) {
        std::vector<uint32_t> split{sc};
        split.insert(split.end(), c.begin(), c.end());
        out.push_back(split);
      }
    }
  }
  return out;
}

struct Timer {
  hipStream_t stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit Timer(hipStream_t s) : stream(s) {}
  ~Timer() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  int init() {
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    return (int)err;
  }
  // mean ms per run of `ls` over enough repeats to fill ~1 ms, best of two groups
  int time(const std::vector<BaPassLaunch> &ls, float &ms) {
    int rc = ba_run_launches(ls, stream);  // warm-up, and a first estimate
    if (rc) return rc;
    hipError_t err = hipEventRecord(e0, stream);
    if (err == hipSuccess) rc = ba_run_launches(ls, stream);
    if (err == hipSuccess && rc == 0) err = hipEventRecord(e1, stream);
    if (err == hipSuccess && rc == 0) err = hipEventSynchronize(e1);
    float one = 0;
    if (err == hipSuccess && rc == 0) err = hipEventElapsedTime(&one, e0, e1);
    if (err != hipSuccess || rc) return rc ? rc : (int)err;
    const int reps = std::max(2, std::min(50, (int)(1.0f / std::max(one, 1e-3f))));
    float best = 1e30f;
    for (int g = 0; g < 2; g++) {
      err = hipEventRecord(e0, stream);
      for (int i = 0; err == hipSuccess && rc == 0 && i < reps; i++) rc = ba_run_launches(ls, stream);
      if (err == hipSuccess && rc == 0) err = hipEventRecord(e1, stream);
      if (err == hipSuccess && rc == 0) err = hipEventSynchronize(e1);
      float t = 0;
      if (err == hipSuccess && rc == 0) err = hipEventElapsedTime(&t, e0, e1);
      if (err != hipSuccess || rc) return rc ? rc : (int)err;
      best = std::min(best, t / reps);
    }
    ms = best;
    return BA_SUCCESS;
  }
};

// kind: 0 plain b2n forward, 1 plain n2b forward, BA_TUNING_LDE_COSETS the LDE's coset NTT (n2b
// forward with coset scale, reading its input bit-reversed through the gathered first pass)
int time_entry(ba_context *ctx, const BaTuned &entry, uint32_t log_n, int kind, const uint64_t *in, uint64_t *out,
               uint32_t cols, Timer &timer, float &ms) {
  ctx->tuned[kind][log_n] = entry;
  const uint64_t n = 1ull << log_n;
  std::vector<BaPassLaunch> ls;
  const bool lde = kind == BA_TUNING_LDE_COSETS;
  const int rc = lde ? ba_build_ntt(ctx, in, out, log_n, cols, n, n, false, false, 1, 1, true, ls, BA_LAT_NEVER)
                     : ba_build_ntt(ctx, in, out, log_n, cols, n, n, kind == 0, false, 0, 0, false, ls, BA_LAT_AUTO);
  if (rc) return rc;
  return timer.time(ls, ms);
}

// the gathered first pass (top strided pass) of an LDE coset NTT exists for 4..8 bits
bool lde_split(const std::vector<uint32_t> &split) {
  for (size_t i = 1; i < split.size(); i++)
    if (split[i] > 8) return false;
  return true;
}

int tune_class(ba_context *ctx, uint32_t log_n, int kind, const uint64_t *in, uint64_t *out, uint64_t elems,
               Timer &timer) {
  const bool n2b = kind != 0;
  const bool lde = kind == BA_TUNING_LDE_COSETS;
  const uint32_t cols = (uint32_t)std::max<uint64_t>(1, elems >> log_n);
  BaTuned best{};
  float best_ms = 1e30f;
  // the 4096-element-tile candidates are timed as such (tile_log 12) whatever the trial size
  for (const auto &split : candidate_splits(log_n)) {
    if (lde && !lde_split(split)) continue;
    BaTuned e{};
    e.tile_log = 12;
    e.valid = ba_plan_from_bits(log_n, n2b, split.data(), (int)split.size(), e.plan);
    if (!e.valid) return BA_ERROR_INVALID_VALUE;  // the enumeration only yields valid splits
    float ms = 0;
    const int rc = time_entry(ctx, e, log_n, kind, in, out, cols, timer, ms);
    if (rc) return rc;
    if (ms < best_ms) best_ms = ms, best = e;
  }
  // then the columns per workgroup of each pass in turn (0 = built-in choice, already timed)
  for (int p = 0; p < best.plan.num_passes; p++) {
    for (uint32_t r : {1u, 2u, 4u, 8u}) {
      BaTuned e = best;
      e.reps[p] = r;
      float ms = 0;
      const int rc = time_entry(ctx, e, log_n, kind, in, out, cols, timer, ms);
      if (rc) return rc;
      if (ms < best_ms * 0.99f) best_ms = ms, best = e;
    }
  }
  // then the waves per SIMD of a 9..12-bit contiguous pass (3 with the next-tile prefetch, or 4)
  bool has_hi_contig = false;
  for (int p = 0; p < best.plan.num_passes; p++) has_hi_contig |= best.plan.contig[p] && best.plan.bits[p] >= 9;
  if (has_hi_contig) {
    for (uint32_t w : {3u, 4u}) {
      BaTuned e = best;
      e.contig_waves = w;
      float ms = 0;
      const int rc = time_entry(ctx, e, log_n, kind, in, out, cols, timer, ms);
      if (rc) return rc;
      if (ms < best_ms * 0.99f) best_ms = ms, best = e;
    }
  }
  // then the latency-shaped passes (1024-element tiles, 4 elements per thread) for every batch
  // size of the class; if they lose, the class keeps the built-in rule (they still run batches of
  // at most 2^20 elements, where they measured faster throughout)
  bool lat = false;
  if (!lde && ba_latency_shape_exists(log_n)) {
    BaTuned e = best;
    e.tile_log = gl::ntt::LAT_TL;
    float ms = 0;
    const int rc = time_entry(ctx, e, log_n, kind, in, out, cols, timer, ms);
    if (rc) return rc;
    lat = ms < best_ms * 0.99f;
  }
  best.tile_log = lat ? gl::ntt::LAT_TL : 0;
  ctx->tuned[kind][log_n] = best;
  return BA_SUCCESS;
}

bool context_ok(ba_context_t ctx) {
  int device = -1;
  return ctx && hipGetDevice(&device) == hipSuccess && device == ctx->device;
}

}  // namespace

extern "C" int ba_autotune(ba_context_t ctx, uint32_t log_n_min, uint32_t log_n_max, uint64_t elems_per_trial,
                           ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (!context_ok(ctx) || log_n_min < 1 || log_n_min > log_n_max || log_n_max > BA_MAX_LOG_N)
      return BA_ERROR_INVALID_VALUE;
    const uint64_t elems = std::max<uint64_t>(elems_per_trial, 1ull << log_n_max);
    uint64_t *in = nullptr, *out = nullptr;
    hipError_t err = hipMalloc(&in, elems * 8);
    if (err == hipSuccess) err = hipMalloc(&out, elems * 8);
    int rc = (int)err;
    const auto saved = std::vector<BaTuned>(&ctx->tuned[0][0], &ctx->tuned[0][0] + 3 * (BA_MAX_LOG_N + 1));
    if (rc == BA_SUCCESS) rc = ba_fill_uniform(in, elems, elems, 1, 0x7475u, stream);
    {
      Timer timer((hipStream_t)stream);
      if (rc == BA_SUCCESS) rc = timer.init();
      for (uint32_t log_n = log_n_min; rc == BA_SUCCESS && log_n <= log_n_max; log_n++)
        for (int kind : {1, 0, BA_TUNING_LDE_COSETS})
          if (rc == BA_SUCCESS) rc = tune_class(ctx, log_n, kind, in, out, elems, timer);
    }
    if (rc != BA_SUCCESS) std::copy(saved.begin(), saved.end(), &ctx->tuned[0][0]);  // keep the old state
    (void)hipStreamSynchronize((hipStream_t)stream);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    return rc;
  });
}

extern "C" int ba_tuning_get(ba_context_t ctx, uint32_t log_n, int bitrev_inputs, ba_tuned_plan *plan) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (!plan || log_n < 1 || log_n > BA_MAX_LOG_N || bitrev_inputs < 0 || bitrev_inputs > BA_TUNING_LDE_COSETS)
      return BA_ERROR_INVALID_VALUE;
    const int kind = bitrev_inputs == BA_TUNING_LDE_COSETS ? BA_TUNING_LDE_COSETS : (bitrev_inputs ? 0 : 1);
    const bool n2b = kind != 0;
    const BaTuned &t = ctx->tuned[kind][log_n];
    const NttPlan p = t.valid ? t.plan : ba_plan_ntt(log_n, n2b);
    *plan = ba_tuned_plan{};
    plan->num_passes = (uint32_t)p.num_passes;
    for (int i = 0; i < p.num_passes; i++) {
      const int low = n2b ? p.num_passes - 1 - i : i;  // execution order -> low-bits-up order
      plan->bits[low] = p.bits[i];
      plan->reps[low] = t.valid ? t.reps[i] : 0;
    }
    plan->contig_waves = t.valid ? t.contig_waves : 0;
    plan->tile_log = t.valid ? t.tile_log : 0;
    return BA_SUCCESS;
  });
}

extern "C" int ba_tuning_set(ba_context_t ctx, uint32_t log_n, int bitrev_inputs, const ba_tuned_plan *plan) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (log_n < 1 || log_n > BA_MAX_LOG_N || bitrev_inputs < 0 || bitrev_inputs > BA_TUNING_LDE_COSETS)
      return BA_ERROR_INVALID_VALUE;
    const int kind = bitrev_inputs == BA_TUNING_LDE_COSETS ? BA_TUNING_LDE_COSETS : (bitrev_inputs ? 0 : 1);
    const bool n2b = kind != 0;
    if (!plan) {
      ctx->tuned[kind][log_n] = BaTuned{};
      return BA_SUCCESS;
    }
    if (plan->num_passes < 1 || plan->num_passes > 8) return BA_ERROR_INVALID_VALUE;
    BaTuned e{};
    if (!ba_plan_from_bits(log_n, n2b, plan->bits, (int)plan->num_passes, e.plan)) return BA_ERROR_INVALID_VALUE;
    for (uint32_t i = 0; i < plan->num_passes; i++) {
      if (plan->reps[i] > 8) return BA_ERROR_INVALID_VALUE;
      const int exec = n2b ? (int)plan->num_passes - 1 - (int)i : (int)i;
      e.reps[exec] = plan->reps[i];
    }
    if (plan->contig_waves != 0 && plan->contig_waves != 3 && plan->contig_waves != 4) return BA_ERROR_INVALID_VALUE;
    e.contig_waves = plan->contig_waves;
    if (plan->tile_log != 0 && plan->tile_log != 12 &&
        !(plan->tile_log == (uint32_t)gl::ntt::LAT_TL && ba_latency_shape_exists(log_n)))
      return BA_ERROR_INVALID_VALUE;
    e.tile_log = plan->tile_log;
    if (kind == BA_TUNING_LDE_COSETS) {  // no latency-shaped tiles, no gathered pass above 8 bits
      if (e.tile_log) return BA_ERROR_INVALID_VALUE;
      for (uint32_t i = 1; i < plan->num_passes; i++)
        if (plan->bits[i] > 8) return BA_ERROR_INVALID_VALUE;
    }
    e.valid = true;
    ctx->tuned[kind][log_n] = e;
    return BA_SUCCESS;
  });
}

extern "C" int ba_tuning_candidates(uint32_t log_n, ba_tuned_plan *plans, uint32_t capacity, uint32_t *count) {
  return ba_guarded([&]() -> int {
    if (!count || log_n < 1 || log_n > BA_MAX_LOG_N || (capacity && !plans)) return BA_ERROR_INVALID_VALUE;
    const auto splits = candidate_splits(log_n);
    *count = (uint32_t)splits.size();
    for (uint32_t i = 0; i < capacity && i < splits.size(); i++) {
      plans[i] = ba_tuned_plan{};
      plans[i].num_passes = (uint32_t)splits[i].size();
      for (size_t k = 0; k < splits[i].size(); k++) plans[i].bits[k] = splits[i][k];
    }
    return BA_SUCCESS;
  });
}
