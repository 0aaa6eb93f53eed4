// ntt_inst.h -- launch-side handles to the pass kernel instantiations, which are spread over
// several translation units (ntt_contig_lo.hip, ntt_contig_hi.hip, ntt_strided.hip) so they
// compile in parallel.
#pragma once

#include "ntt_kernels.h"

typedef void (*BaPassKernel)(gl::ntt::PassArgs);

// contiguous pass over bits [0, S), 4 <= S <= 8 / 9 <= S <= 12
BaPassKernel ba_pick_contig_lo(uint32_t S, bool n2b, bool inverse, bool scale, bool whole);
BaPassKernel ba_pick_contig_hi(uint32_t S, bool n2b, bool inverse, bool scale, bool whole);
// strided pass, 4 <= S <= 8; top: the pass holds the top bits; scale: fused (pre/post) scale;
// gather: the n2b top pass reads bit-reversed columns
BaPassKernel ba_pick_strided(uint32_t S, bool n2b, bool inverse, bool top, bool scale, bool gather);
// whole transform of log_n <= 3 per thread
BaPassKernel ba_pick_tiny(uint32_t log_n, bool n2b, bool inverse);
