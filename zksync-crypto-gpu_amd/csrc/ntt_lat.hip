// ntt_lat.hip -- instantiations of the latency-shaped NTT passes (see ntt_lat.h) and their
// kernel names for launch tracing.
#include <cstdio>

#include "ntt_inst.h"
#include "ntt_lat.h"

using namespace gl::ntt;

template <int S, bool N2B, bool INV, bool STRIDED, bool TOP, bool SCALE>
static BaKernels one() {
  return {ntt_pass_lat<S, N2B, INV, STRIDED, TOP, SCALE>, ntt_pass_lat_group<S, N2B, INV, STRIDED, TOP, SCALE>};
}

template <int S, bool N2B, bool INV>
static BaKernels pick3(bool strided, bool top, bool scale) {
  if (!strided) {
    if constexpr (S >= BA_LAT_MIN_CONTIG && S <= BA_LAT_MAX_CONTIG)
      return scale ? one<S, N2B, INV, false, false, true>() : one<S, N2B, INV, false, false, false>();
  } else {
    if constexpr (S >= BA_LAT_MIN_STRIDED && S <= BA_LAT_MAX_STRIDED) {
      if (!top) return scale ? BaKernels{nullptr, nullptr} : one<S, N2B, INV, true, false, false>();
      return scale ? one<S, N2B, INV, true, true, true>() : one<S, N2B, INV, true, true, false>();
    }
  }
  return {nullptr, nullptr};
}

template <int S>
static BaKernels pick(bool strided, bool n2b, bool inverse, bool top, bool scale) {
  if (n2b) return inverse ? pick3<S, true, true>(strided, top, scale) : pick3<S, true, false>(strided, top, scale);
  return inverse ? pick3<S, false, true>(strided, top, scale) : pick3<S, false, false>(strided, top, scale);
}

BaKernels ba_pick_lat(uint32_t S, bool strided, bool n2b, bool inverse, bool top, bool scale) {
  switch (S) {
    case 4: return pick<4>(strided, n2b, inverse, top, scale);
    case 5: return pick<5>(strided, n2b, inverse, top, scale);
    case 6: return pick<6>(strided, n2b, inverse, top, scale);
    case 7: return pick<7>(strided, n2b, inverse, top, scale);
    case 8: return pick<8>(strided, n2b, inverse, top, scale);
    case 9: return pick<9>(strided, n2b, inverse, top, scale);
    case 10: return pick<10>(strided, n2b, inverse, top, scale);
  }
  return {nullptr, nullptr};
}

namespace {
// [S][n2b][inverse][strided][top][scale], as rocprofv3 prints them; built by the first caller
// (a function-local static initialiser, so concurrent first calls are safe)
struct LatNames {
  char n[11][2][2][2][2][2][100];
  LatNames() {
    static const char *tf[2] = {"false", "true"};
    for (int s = 0; s < 11; s++)
      for (int d = 0; d < 2; d++)
        for (int i = 0; i < 2; i++)
          for (int st = 0; st < 2; st++)
            for (int tp = 0; tp < 2; tp++)
              for (int sc = 0; sc < 2; sc++)
                snprintf(n[s][d][i][st][tp][sc], sizeof(n[s][d][i][st][tp][sc]),
                         "void gl::ntt::ntt_pass_lat<%d, %s, %s, %s, %s, %s>(gl::ntt::PassArgs)", s, tf[d], tf[i],
                         tf[st], tf[tp], tf[sc]);
  }
};
}  // namespace

const char *ba_lat_kernel_name(uint32_t S, bool strided, bool n2b, bool inverse, bool top, bool scale) {
  static const LatNames names;
  return S <= 10 ? names.n[S][n2b][inverse][strided][top][scale] : "ntt_pass_lat";
}
