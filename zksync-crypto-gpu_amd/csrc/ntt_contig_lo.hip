// ntt_contig_lo.hip -- instantiations of the contiguous NTT pass (see ntt_kernels.h).
#include "ntt_inst.h"

using namespace gl::ntt;

// short columns (a tile holds whole columns) and 4..8-bit passes: the three-waves variant
template <int S, bool N2B, bool INV, int SC, bool WHOLE>
static BaKernels pair() {
  return {ntt_pass_contig<S, N2B, INV, SC, WHOLE, 3>, ntt_pass_contig_group<S, N2B, INV, SC, WHOLE, 3>};
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

BaKernels ba_pick_contig_lo(uint32_t S, bool n2b, bool inverse, int scale, bool whole) {
  switch (S) {
    case 4: return pick<4>(n2b, inverse, scale, whole);
    case 5: return pick<5>(n2b, inverse, scale, whole);
    case 6: return pick<6>(n2b, inverse, scale, whole);
    case 7: return pick<7>(n2b, inverse, scale, whole);
    case 8: return pick<8>(n2b, inverse, scale, whole);
  }
  return {nullptr, nullptr};
}
