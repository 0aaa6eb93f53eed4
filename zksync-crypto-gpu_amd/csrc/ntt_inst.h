// ntt_inst.h -- launch-side handles to the pass kernel instantiations, which are spread over
// several translation units (ntt_contig_lo.hip, ntt_contig_hi.hip, ntt_strided.hip) so they
// compile in parallel.  Every pass shape has a single-batch kernel and a grouped kernel.
#pragma once

#include "ntt_kernels.h"

typedef void (*BaPassKernel)(gl::ntt::PassArgs);
typedef void (*BaGroupKernel)(const gl::ntt::PassArgs *, const uint32_t *, uint32_t);

struct BaKernels {
  BaPassKernel single;
  BaGroupKernel group;
  bool operator==(const BaKernels &o) const { return single == o.single; }
  bool operator<(const BaKernels &o) const { return (void *)single < (void *)o.single; }
};

// contiguous pass over bits [0, S), 4 <= S <= 8 / 9 <= S <= 12; scale: 0 none, 1 one factor
// for every element, 2 per-element coset factors
BaKernels ba_pick_contig_lo(uint32_t S, bool n2b, bool inverse, int scale, bool whole);
BaKernels ba_pick_contig_hi(uint32_t S, bool n2b, bool inverse, int scale, bool whole);   // 4 waves/SIMD
BaKernels ba_pick_contig_hi3(uint32_t S, bool n2b, bool inverse, int scale, bool whole);  // 3, prefetch
// strided pass, 4 <= S <= 10; top: the pass holds the top bits; scale: fused (pre/post) scale;
// gather: the n2b top pass reads bit-reversed columns
BaKernels ba_pick_strided(uint32_t S, bool n2b, bool inverse, bool top, bool scale, bool gather, bool shared);
// whole transform of log_n <= 3 per thread
BaKernels ba_pick_tiny(uint32_t log_n, bool n2b, bool inverse);
// latency-shaped passes (ntt_lat.h), 1024-element tiles: contiguous over 4..10 bits, strided over
// 5..8 bits
constexpr int BA_LAT_MIN_CONTIG = 4, BA_LAT_MAX_CONTIG = 10;
constexpr int BA_LAT_MIN_STRIDED = 5, BA_LAT_MAX_STRIDED = 8;
BaKernels ba_pick_lat(uint32_t S, bool strided, bool n2b, bool inverse, bool top, bool scale);
const char *ba_lat_kernel_name(uint32_t S, bool strided, bool n2b, bool inverse, bool top, bool scale);
