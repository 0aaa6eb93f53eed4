// ntt_contig_hi.hip -- instantiations of the contiguous NTT pass (see ntt_kernels.h) over 9..12
// bits, at BA_CONTIG_W waves per SIMD (this file is compiled once per variant: 4, and as
// ntt_contig_hi3.hip, 3).
#include "ntt_inst.h"

#ifndef BA_CONTIG_W
#define BA_CONTIG_W 4
#endif

using namespace gl::ntt;

template <int S, bool N2B, bool INV, int SC, bool WHOLE>
static BaKernels pair() {
  return {ntt_pass_contig<S, N2B, INV, SC, WHOLE, BA_CONTIG_W>, ntt_pass_contig_group<S, N2B, INV, SC, WHOLE, BA_CONTIG_W>};
}

// columns of >= 4096 elements use S = 12, or S = 9..11 when log_n is 13..15; shorter
// columns are transformed in one pass with S = log_n
template <int S, bool N2B, bool INV, int SC>
static BaKernels pick4(bool whole) {
  if constexpr (S == 12) return pair<S, N2B, INV, SC, true>();
  else if constexpr (S < 9) return pair<S, N2B, INV, SC, false>();
  else return whole ? pair<S, N2B, INV, SC, true>() : pair<S, N2B, INV, SC, false>();
}

// scale: 0 none, 1 one factor for all elements, 2 per-element coset factors
template <int S, bool N2B, bool INV>
static BaKernels pick_sc(int scale, bool whole) {
  if (scale == 2) return pick4<S, N2B, INV, 2>(whole);
  if (scale == 1) return pick4<S, N2B, INV, 1>(whole);
  return pick4<S, N2B, INV, 0>(whole);
}

template <int S>
static BaKernels pick(bool n2b, bool inverse, int scale, bool whole) {
  if (n2b) return inverse ? pick_sc<S, true, true>(scale, whole) : pick_sc<S, true, false>(scale, whole);
  return inverse ? pick_sc<S, false, true>(scale, whole) : pick_sc<S, false, false>(scale, whole);
}

#if BA_CONTIG_W == 3
BaKernels ba_pick_contig_hi3(uint32_t S, bool n2b, bool inverse, int scale, bool whole) {
#else
BaKernels ba_pick_contig_hi(uint32_t S, bool n2b, bool inverse, int scale, bool whole) {
#endif
  switch (S) {
    case 9: return pick<9>(n2b, inverse, scale, whole);
    case 10: return pick<10>(n2b, inverse, scale, whole);
    case 11: return pick<11>(n2b, inverse, scale, whole);
    case 12: return pick<12>(n2b, inverse, scale, whole);
  }
  return {nullptr, nullptr};
}
