// ntt_strided.hip -- instantiations of the strided NTT pass and the tiny transform (see
// ntt_kernels.h).
#include "ntt_inst.h"

using namespace gl::ntt;

// shared: the launch shares the GPU with other streams' work; a four-waves variant then comes in
// its three-waves build (ntt_pass_strided_shared), the others are the same either way
template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER>
static BaKernels pair(bool shared) {
  if constexpr (strided_four_waves<S, TOP, GATHER>()) {
    if (shared)
      return {ntt_pass_strided_shared<S, N2B, INV, TOP, SCALE, GATHER>,
              ntt_pass_strided_shared_group<S, N2B, INV, TOP, SCALE, GATHER>};
  }
  return {ntt_pass_strided<S, N2B, INV, TOP, SCALE, GATHER>, ntt_pass_strided_group<S, N2B, INV, TOP, SCALE, GATHER>};
}

// the gathered top pass (LDE coset NTTs) exists for 4..8 bits only
template <int S, bool N2B, bool INV>
static BaKernels pick3(bool top, bool scale, bool gather, bool shared) {
  if (!top) return gather ? BaKernels{nullptr, nullptr} : pair<S, N2B, INV, false, false, false>(shared);
  if (gather) {
    if constexpr (N2B && S <= 8)
      return scale ? pair<S, N2B, INV, true, true, true>(shared) : pair<S, N2B, INV, true, false, true>(shared);
    return {nullptr, nullptr};
  }
  return scale ? pair<S, N2B, INV, true, true, false>(shared) : pair<S, N2B, INV, true, false, false>(shared);
}

template <int S>
static BaKernels pick(bool n2b, bool inverse, bool top, bool scale, bool gather, bool shared) {
  if (n2b)
    return inverse ? pick3<S, true, true>(top, scale, gather, shared) : pick3<S, true, false>(top, scale, gather, shared);
  return inverse ? pick3<S, false, true>(top, scale, gather, shared) : pick3<S, false, false>(top, scale, gather, shared);
}

BaKernels ba_pick_strided(uint32_t S, bool n2b, bool inverse, bool top, bool scale, bool gather, bool shared) {
  switch (S) {
    case 4: return pick<4>(n2b, inverse, top, scale, gather, shared);
    case 5: return pick<5>(n2b, inverse, top, scale, gather, shared);
    case 6: return pick<6>(n2b, inverse, top, scale, gather, shared);
    case 7: return pick<7>(n2b, inverse, top, scale, gather, shared);
    case 8: return pick<8>(n2b, inverse, top, scale, gather, shared);
    case 9: return pick<9>(n2b, inverse, top, scale, gather, shared);
    case 10: return pick<10>(n2b, inverse, top, scale, gather, shared);
  }
  return {nullptr, nullptr};
}

template <int K, bool N2B, bool INV>
static BaKernels tiny() {
  return {ntt_tiny<K, N2B, INV>, ntt_tiny_group<K, N2B, INV>};
}

BaKernels ba_pick_tiny(uint32_t log_n, bool n2b, bool inverse) {
#define BA_TINY(K)                                                              \
  if (log_n == K) {                                                             \
    if (n2b) return inverse ? tiny<K, true, true>() : tiny<K, true, false>();   \
    return inverse ? tiny<K, false, true>() : tiny<K, false, false>();          \
  }
  BA_TINY(1)
  BA_TINY(2)
  BA_TINY(3)
#undef BA_TINY
  return {nullptr, nullptr};
}
