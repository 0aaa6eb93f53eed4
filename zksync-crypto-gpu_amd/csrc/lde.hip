// lde.hip -- batched coset low-degree extension (shivini's trace-commitment LDE).
//
// Reference composition (crates/shivini/src/data_structures/storage.rs:154-260,
// crates/shivini/src/primitives/ntt.rs:10-42,407-510): evaluations -> n2b inverse NTT -> bit
// reverse -> monomials (natural) -> for each coset c < B: n2b forward NTT with coset index
// bitrev_{log B}(c) and log_extension_degree = log B.
//
// Here the monomials stay in the bit-reversed order the inverse NTT leaves them in, and the
// first pass of every coset NTT gathers them (coalesced chunk reads staged through LDS), so the
// separate bit-reverse pass over the data disappears.  Launches: passes(iNTT) + B * passes(NTT),
// all on one stream, capturable into a hipGraph.
#include "api_internal.h"

namespace {

uint32_t bitrev_bits(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

}  // namespace

extern "C" int ba_lde_batch(ba_context_t ctx, uint64_t *evals, uint64_t in_stride,
                            uint64_t *monomials, uint64_t *out, uint64_t out_col_stride,
                            uint64_t out_coset_stride, uint32_t log_n, uint32_t num_cols,
                            uint32_t log_blowup, ba_stream_t stream) {
  if (!ctx) return BA_ERROR_NOT_INITIALIZED;
  if (log_n < 1 || log_n > BA_MAX_LOG_N || log_blowup < 1 || log_n + log_blowup > gl::OMEGA_LOG_ORDER)
    return BA_ERROR_INVALID_VALUE;
  const uint64_t n = 1ull << log_n;
  const uint32_t blowup = 1u << log_blowup;
  if (num_cols == 0) return BA_SUCCESS;
  if (!evals || !out || in_stride < n) return BA_ERROR_INVALID_VALUE;
  // cosets of one column must not overlap each other or the next column
  if (out_coset_stride < n && blowup > 1) return BA_ERROR_INVALID_VALUE;
  if (num_cols > 1 && out_col_stride < n) return BA_ERROR_INVALID_VALUE;
  if ((void *)out == (void *)evals || (monomials && (void *)out == (void *)monomials))
    return BA_ERROR_INVALID_VALUE;
  hipStream_t s = (hipStream_t)stream;
  uint64_t *mono = monomials ? monomials : evals;
  const uint64_t mono_stride = monomials ? n : in_stride;
  // monomials, bit-reversed: n2b inverse without coset factors
  int rc = ba_launch_ntt(ctx, evals, mono, log_n, num_cols, in_stride, mono_stride, false, true, 0, 0, s);
  if (rc) return rc;
  for (uint32_t c = 0; c < blowup; c++) {
    rc = ba_launch_ntt(ctx, mono, out + (uint64_t)c * out_coset_stride, log_n, num_cols, mono_stride,
                       out_col_stride, false, false, log_blowup, bitrev_bits(c, log_blowup), s,
                       /*in_bitrev_gather=*/true);
    if (rc) return rc;
  }
  return BA_SUCCESS;
}
