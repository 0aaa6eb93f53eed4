// ntt_contig_hi3.hip -- the three-waves-per-SIMD (next-tile prefetch) variant of the 9..12-bit
// contiguous pass instantiations, compiled in parallel with ntt_contig_hi.hip.
#define BA_CONTIG_W 3
#include "ntt_contig_hi.hip"
