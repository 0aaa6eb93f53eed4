// ntt_strided.hip -- instantiations of the strided NTT pass and the tiny transform (see
// ntt_kernels.h).
#include "ntt_inst.h"

using namespace gl::ntt;

template <int S, bool N2B, bool INV>
static BaPassKernel pick3(bool top, bool scale, bool gather) {
  if (!top) return ntt_pass_strided<S, N2B, INV, false, false, false>;
  if constexpr (N2B) {
    if (gather)
      return scale ? ntt_pass_strided<S, N2B, INV, true, true, true>
                   : ntt_pass_strided<S, N2B, INV, true, false, true>;
  }
  return scale ? ntt_pass_strided<S, N2B, INV, true, true, false>
               : ntt_pass_strided<S, N2B, INV, true, false, false>;
}

template <int S>
static BaPassKernel pick(bool n2b, bool inverse, bool top, bool scale, bool gather) {
  if (n2b) return inverse ? pick3<S, true, true>(top, scale, gather) : pick3<S, true, false>(top, scale, gather);
  return inverse ? pick3<S, false, true>(top, scale, gather) : pick3<S, false, false>(top, scale, gather);
}

BaPassKernel ba_pick_strided(uint32_t S, bool n2b, bool inverse, bool top, bool scale, bool gather) {
  switch (S) {
    case 4: return pick<4>(n2b, inverse, top, scale, gather);
    case 5: return pick<5>(n2b, inverse, top, scale, gather);
    case 6: return pick<6>(n2b, inverse, top, scale, gather);
    case 7: return pick<7>(n2b, inverse, top, scale, gather);
    case 8: return pick<8>(n2b, inverse, top, scale, gather);
  }
  return nullptr;
}

BaPassKernel ba_pick_tiny(uint32_t log_n, bool n2b, bool inverse) {
#define BA_TINY(K)                                                                \
  if (log_n == K) {                                                               \
    if (n2b) return inverse ? ntt_tiny<K, true, true> : ntt_tiny<K, true, false>; \
    return inverse ? ntt_tiny<K, false, true> : ntt_tiny<K, false, false>;        \
  }
  BA_TINY(1)
  BA_TINY(2)
  BA_TINY(3)
#undef BA_TINY
  return nullptr;
}
