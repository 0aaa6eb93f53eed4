// ntt_contig_hi.hip -- instantiations of the contiguous NTT pass (see ntt_kernels.h).
#include "ntt_inst.h"

using namespace gl::ntt;

template <int S, bool N2B, bool INV, bool SCALE, bool WHOLE>
static BaKernels pair() {
  return {ntt_pass_contig<S, N2B, INV, SCALE, WHOLE>, ntt_pass_contig_group<S, N2B, INV, SCALE, WHOLE>};
}

// columns of >= 4096 elements use S = 12, or S = 9..11 when log_n is 13..15; shorter
// columns are transformed in one pass with S = log_n
template <int S, bool N2B, bool INV, bool SCALE>
static BaKernels pick4(bool whole) {
  if constexpr (S == 12) return pair<S, N2B, INV, SCALE, true>();
  else if constexpr (S < 9) return pair<S, N2B, INV, SCALE, false>();
  else return whole ? pair<S, N2B, INV, SCALE, true>() : pair<S, N2B, INV, SCALE, false>();
}

template <int S>
static BaKernels pick(bool n2b, bool inverse, bool scale, bool whole) {
  if (n2b) {
    if (inverse) return scale ? pick4<S, true, true, true>(whole) : pick4<S, true, true, false>(whole);
    return scale ? pick4<S, true, false, true>(whole) : pick4<S, true, false, false>(whole);
  }
  if (inverse) return scale ? pick4<S, false, true, true>(whole) : pick4<S, false, true, false>(whole);
  return scale ? pick4<S, false, false, true>(whole) : pick4<S, false, false, false>(whole);
}

BaKernels ba_pick_contig_hi(uint32_t S, bool n2b, bool inverse, bool scale, bool whole) {
  switch (S) {
    case 9: return pick<9>(n2b, inverse, scale, whole);
    case 10: return pick<10>(n2b, inverse, scale, whole);
    case 11: return pick<11>(n2b, inverse, scale, whole);
    case 12: return pick<12>(n2b, inverse, scale, whole);
  }
  return {nullptr, nullptr};
}
