// lde.hip -- batched coset low-degree extension (shivini's trace-commitment LDE).
//
// Reference composition (crates/shivini/src/data_structures/storage.rs:154-260,
// crates/shivini/src/primitives/ntt.rs:10-42,407-510): evaluations -> n2b inverse NTT -> bit
// reverse -> monomials (natural) -> for each coset c < B: n2b forward NTT with coset index
// bitrev_{log B}(c) and log_extension_degree = log B.
//
// Here the monomials stay in the bit-reversed order the inverse NTT leaves them in, and the
// first pass of every coset NTT gathers them (coalesced chunk reads staged through LDS), so the
// separate bit-reverse pass over the data disappears.  Launches: passes(iNTT) + B * passes(NTT);
// the inverse NTT and coset 0 on the caller's stream, cosets 1.. on the context's auxiliary
// streams forked from it and joined back, capturable into a hipGraph.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "api_internal.h"

namespace {

uint32_t bitrev_bits(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// True when the n-element vectors starting at k s1 + c s2 (k < K, c < C) are pairwise disjoint.
// Two starts differ by dk s1 + dc s2 with |dk| < K, |dc| < C, not both zero.  Up to sign the
// index with the smaller count (outer) is >= 0; for each outer step the point of the inner
// progression closest to it is found by division, so no list of starts is built and the work
// is O(min(K, C)).
bool vectors_disjoint(uint64_t K, uint64_t s1, uint64_t C, uint64_t s2, uint64_t n) {
  typedef unsigned __int128 u128;
  uint64_t O = K, so = s1, I = C, si = s2;
  if (C < K) O = C, so = s2, I = K, si = s1;
  if (I > 1 && si < n) return false;  // outer step 0: neighbours along the inner stride
  for (uint64_t d = 1; d < O; d++) {
    const u128 x = (u128)d * so;  // distance to cancel with inner steps -e, 0 <= e <= I - 1
    u128 closest;
    if (si == 0) {
      closest = x;
    } else {
      const u128 q = x / si, r = x % si;
      if (q > I - 1) {
        closest = x - (u128)(I - 1) * si;
      } else {
        closest = r;
        if (q + 1 <= I - 1 && (u128)si - r < closest) closest = (u128)si - r;
      }
    }
    if (closest < n) return false;
  }
  return true;
}

}  // namespace

int ba_validate_lde(uint64_t *evals, uint64_t in_stride, uint64_t *monomials, uint64_t *out,
                    uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                    uint32_t num_cols, uint32_t log_blowup) {
  if (log_n < 1 || log_n > BA_MAX_LOG_N || log_blowup < 1 || log_n + log_blowup > gl::OMEGA_LOG_ORDER)
    return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  const uint64_t blowup = 1ull << log_blowup;
  if (num_cols == 0) return BA_SUCCESS;
  if (!evals || !out || in_stride < n) return BA_ERROR_INVALID_VALUE;
  // every buffer's extent must be addressable
  typedef unsigned __int128 u128;
  const u128 out_span = ((u128)(num_cols - 1) * out_col_stride + (u128)(blowup - 1) * out_coset_stride + n) * 8;
  const u128 ev_span = ((u128)(num_cols - 1) * in_stride + n) * 8;
  const u128 mono_span = (u128)num_cols * n * 8;
  if ((u128)(uintptr_t)out + out_span > ((u128)1 << 64) || (u128)(uintptr_t)evals + ev_span > ((u128)1 << 64) ||
      (monomials && (u128)(uintptr_t)monomials + mono_span > ((u128)1 << 64)))
    return BA_ERROR_INVALID_VALUE;
  // the B * num_cols output vectors of n elements must be pairwise disjoint
  if (!vectors_disjoint(num_cols, out_col_stride, blowup, out_coset_stride, n)) return BA_ERROR_INVALID_VALUE;
  // the outputs overlap neither the evaluations nor the monomials; monomials given separately
  // do not overlap the evaluations
  const char *out_end = (const char *)out + (size_t)out_span;
  const char *ev_end = (const char *)evals + (size_t)ev_span;
  auto overlap = [](const void *a0, const void *a1, const void *b0, const void *b1) {
    return a0 < b1 && b0 < a1;
  };
  if (overlap(out, out_end, evals, ev_end)) return BA_ERROR_INVALID_VALUE;
  if (monomials) {
    const char *mono_end = (const char *)monomials + (size_t)mono_span;
    if (overlap(out, out_end, monomials, mono_end) || overlap(evals, ev_end, monomials, mono_end))
      return BA_ERROR_INVALID_VALUE;
  }
  return BA_SUCCESS;
}

int ba_build_lde(const ba_context *ctx, uint64_t *evals, uint64_t in_stride, uint64_t *monomials,
                 uint64_t *out, uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                 uint32_t num_cols, uint32_t log_blowup, std::vector<BaLeveledLaunch> &launches,
                 uint32_t first_level, size_t *inverse_launches) {
  if (num_cols == 0) return BA_SUCCESS;
  const uint64_t n = 1ull << log_n;
  const uint32_t blowup = 1u << log_blowup;
  uint64_t *mono = monomials ? monomials : evals;
  const uint64_t mono_stride = monomials ? n : in_stride;
  // monomials, bit-reversed: n2b inverse without coset factors
  std::vector<BaPassLaunch> ls;
  int rc = ba_build_ntt(ctx, evals, mono, log_n, num_cols, in_stride, mono_stride, false, true, 0, 0, false, ls);
  if (rc) return rc;
  uint32_t depth = 0;
  for (const BaPassLaunch &l : ls) {
    launches.push_back({first_level + (uint32_t)l.pass, l});
    depth = std::max<uint32_t>(depth, (uint32_t)l.pass + 1);
  }
  if (inverse_launches) *inverse_launches = ls.size();
  for (uint32_t c = 0; c < blowup; c++) {
    ls.clear();
    rc = ba_build_ntt(ctx, mono, out + (uint64_t)c * out_coset_stride, log_n, num_cols, mono_stride,
                      out_col_stride, false, false, log_blowup, bitrev_bits(c, log_blowup),
                      /*in_bitrev_gather=*/true, ls);
    if (rc) return rc;
    for (const BaPassLaunch &l : ls) launches.push_back({first_level + depth + (uint32_t)l.pass, l});
  }
  return BA_SUCCESS;
}

// BOOJUM_AMD_LDE_SERIAL=1 in the environment (read once) keeps every LDE's cosets in order on the
// caller's stream, so that a profiler's per-kernel durations are each launch's own.
static bool lde_serial() {
  static const bool serial = [] {
    const char *e = std::getenv("BOOJUM_AMD_LDE_SERIAL");
    return e && e[0] == '1';
  }();
  return serial;
}

// The LDE's launches on `stream` (cosets after the first forked onto the context's auxiliary
// streams and joined back); with natural_monomials the monomials are bit-reversed in place once
// every coset has read them.
static int lde_run(ba_context_t ctx, uint64_t *evals, uint64_t in_stride, uint64_t *monomials,
                   uint64_t *out, uint64_t out_col_stride, uint64_t out_coset_stride, uint32_t log_n,
                   uint32_t num_cols, uint32_t log_blowup, bool natural_monomials, ba_stream_t stream) {
  uint64_t *mono = monomials ? monomials : evals;
  const uint64_t mono_stride = monomials ? (1ull << log_n) : in_stride;
  const int rc = [&]() -> int {
    // arguments first (host-only checks), then the context
    int rc = ba_validate_lde(evals, in_stride, monomials, out, out_col_stride, out_coset_stride, log_n,
                             num_cols, log_blowup);
    if (rc) return rc;
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (num_cols == 0) return BA_SUCCESS;
    int device = -1;  // the context's tables, and its auxiliary streams, live on its device
    if (hipGetDevice(&device) != hipSuccess || device != ctx->device) return BA_ERROR_INVALID_VALUE;
    std::vector<BaLeveledLaunch> leveled;
    size_t n_inv = 0;  // the inverse NTT's launches, first in the list
    rc = ba_build_lde(ctx, evals, in_stride, monomials, out, out_col_stride, out_coset_stride, log_n,
                      num_cols, log_blowup, leveled, 0, &n_inv);
    if (rc) return rc;
    std::vector<BaPassLaunch> flat;
    for (const BaLeveledLaunch &l : leveled) flat.push_back(l.launch);  // already in dependency order
    // flat = the inverse NTT's passes (their own plan), then each coset's (one plan shared by the
    // cosets, which may have another pass count).  The coset NTTs only read the monomials, so
    // cosets 1.. run on auxiliary streams forked after the inverse NTT and joined back: a pass's
    // last partial wave of workgroups then overlaps another coset's passes (a graph captured from
    // `stream` holds the same fork/join).
    const uint32_t blowup = 1u << log_blowup;
    const size_t per = (flat.size() - n_inv) / blowup;
    // (under launch tracing the launches stay in order on one stream, so each one's timing is its own)
    if (blowup == 1 || per == 0 || n_inv == 0 || n_inv + per * blowup != flat.size() || ba_trace_active() ||
        !ctx->lde_fork || lde_serial())
      return ba_run_launches(flat, (hipStream_t)stream);
    const hipStream_t main = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(ctx->lde_mutex);
    hipError_t err = hipSuccess;
    const int aux = (int)std::min<uint32_t>(blowup - 1, ba_context::kLdeAux);
    rc = ba_run_launches(std::vector<BaPassLaunch>(flat.begin(), flat.begin() + n_inv), main);
    if (rc) return rc;
    if ((err = hipEventRecord(ctx->lde_fork, main)) != hipSuccess) return (int)err;
    // Once forked, every auxiliary stream that waited on the fork is joined back even when a later
    // step fails, so that `stream` still orders after all enqueued work and a capture is not left
    // with unjoined streams; the first error is returned.
    int forked = 0;
    for (; forked < aux && !rc; forked++)
      if ((err = hipStreamWaitEvent(ctx->lde_aux[forked], ctx->lde_fork, 0)) != hipSuccess) rc = (int)err;
    if (rc) forked--;  // the failed wait did not fork that stream
    for (uint32_t c = 0; c < blowup && !rc; c++) {
      const hipStream_t sc = c == 0 ? main : ctx->lde_aux[(c - 1) % aux];
      const auto first = flat.begin() + n_inv + per * c;
      rc = ba_run_launches(std::vector<BaPassLaunch>(first, first + per), sc);
    }
    for (int i = 0; i < forked; i++) {
      if ((err = hipEventRecord(ctx->lde_join[i], ctx->lde_aux[i])) == hipSuccess)
        err = hipStreamWaitEvent(main, ctx->lde_join[i], 0);
      if (err != hipSuccess && !rc) rc = (int)err;
    }
    return rc;
  }();
  if (rc || !natural_monomials || num_cols == 0) return rc;
  return ba_bit_reverse_in_place(mono, mono_stride, log_n, num_cols, stream);
}

extern "C" int ba_lde_batch(ba_context_t ctx, uint64_t *evals, uint64_t in_stride,
                            uint64_t *monomials, uint64_t *out, uint64_t out_col_stride,
                            uint64_t out_coset_stride, uint32_t log_n, uint32_t num_cols,
                            uint32_t log_blowup, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    return lde_run(ctx, evals, in_stride, monomials, out, out_col_stride, out_coset_stride, log_n, num_cols,
                   log_blowup, false, stream);
  });
}

extern "C" int ba_lde_batch_ex(ba_context_t ctx, uint64_t *evals, uint64_t in_stride,
                               uint64_t *monomials, uint64_t *out, uint64_t out_col_stride,
                               uint64_t out_coset_stride, uint32_t log_n, uint32_t num_cols,
                               uint32_t log_blowup, uint32_t flags, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    if (flags & ~(uint32_t)BA_LDE_NATURAL_MONOMIALS) return BA_ERROR_INVALID_VALUE;
    return lde_run(ctx, evals, in_stride, monomials, out, out_col_stride, out_coset_stride, log_n, num_cols,
                   log_blowup, (flags & BA_LDE_NATURAL_MONOMIALS) != 0, stream);
  });
}
