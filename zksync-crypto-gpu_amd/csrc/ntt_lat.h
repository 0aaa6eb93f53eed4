// ntt_lat.h -- latency-shaped NTT passes for batches too small to fill the chip.
//
// The throughput passes (ntt_kernels.h) give each thread 16 elements and each workgroup a
// 4096-element tile, so a single 2^16 column is 16 workgroups whose threads each run a ~2900
// instruction dependent chain at one wave per SIMD.  These passes use the same radix-2 flows
// (n2b DIT from the top bits down, b2n DIF from the bottom up, blocks of K consecutive stages
// as x_i *= W^i plus a 2^K-point transform with shift twiddles) in radix-4 blocks: 4 elements
// per thread, a 1024-element tile per 256-thread workgroup, so a 2^16 column is 64 workgroups
// per pass and a thread's chain is a quarter as long.  Twiddle powers W^i of every round are
// read from the context's power tables at the start (independent of the data, so they overlap
// the loads); the first round reads global memory in its own layout and the last round writes
// it, the rounds between exchange through LDS with one barrier each.
//   * contiguous pass: bits [0, S), 2 <= S <= 10; a tile is 2^(10-S) chunks of 2^S (columns
//     of fewer than 1024 elements: several whole columns);
//   * strided pass: bits [P0, P0 + S), S <= 8, P0 >= 10 - S; a tile is 2^S rows of
//     2^(10-S) >= 4 consecutive elements.
// Fused scaling as in the throughput passes: register m of the first (forward pre-scale) or
// last (inverse post-scale) round sits at natural index j0 + pi(m) 2^(L-2), so its factor is
// F(j0) T[4 pi(m)].
#pragma once

#include "ntt_kernels.h"

namespace gl {
namespace ntt {

constexpr int LAT_TL = 10;
constexpr int LAT_TILE = 1 << LAT_TL;
static_assert(LAT_TILE / 4 == THREADS, "a latency tile is 4 elements per thread");

// Round shapes of an S-bit pass in 2-bit windows (the Shape of ntt_kernels.h with 2 for 4):
// n2b from the top bits down (a 1-bit round first when S is odd), b2n from the bottom up (the
// 1-bit round last).  A 1-bit round's window is [WB, WB + 2) with only bit WB + 1 active.
template <int S, bool N2B>
struct Shape2 {
  static_assert(S >= 2, "a latency pass covers at least 2 bits");
  static constexpr int K0 = S & 1;
  static constexpr int FULL = S >> 1;
  static constexpr int ROUNDS = FULL + K0;
  static constexpr int K(int r) {
    if (N2B) return (K0 && r == 0) ? 1 : 2;
    return r < FULL ? 2 : 1;
  }
  static constexpr int WB(int r) {
    if (N2B) return (K0 && r == 0) ? S - 2 : S - K0 - 2 * ((K0 ? r - 1 : r) + 1);
    return r < FULL ? 2 * r : S - 2;
  }
  static constexpr int AB(int r) { return WB(r) + 2 - K(r); }
  static constexpr int TOP = N2B ? 0 : ROUNDS - 1;
};

// local index bits of register m in the round whose 2-bit window starts at bit WB
template <int WB>
__device__ __forceinline__ u32 window2(u32 q, u32 m) {
  return ((q >> WB) << (WB + 2)) | (m << WB) | (q & ((1u << WB) - 1));
}

// STRIDED: bits [lo_bit, lo_bit + S) (TOP: they are the top bits); else bits [0, S).
// SCALE: this pass applies the fused scale (forward: before its first round, which then is
// the transform's first; inverse: after its last).
template <int S, bool N2B, bool INV, bool STRIDED, bool TOP, bool SCALE>
__device__ __forceinline__ void lat_body(const PassArgs &a, const u32 bid) {
  using Sh = Shape2<S, N2B>;
  constexpr int R = Sh::ROUNDS;
  constexpr int LOG_C = LAT_TL - S;
  constexpr int C = 1 << LOG_C;
  static_assert(!STRIDED || LOG_C >= 2, "strided latency passes keep runs of >= 4 elements");
  __shared__ u64 lds[LAT_TILE + LAT_TILE / 16];
  const u32 t = threadIdx.x;
  const u32 L = a.log_n;
  // thread -> (c, q): c picks the run (strided) or chunk (contiguous), q the element pairs;
  // consecutive threads take consecutive addresses in the strided layout and in the
  // contiguous layout's top-bit windows
  const u32 c = STRIDED ? (t & (C - 1)) : (t >> (S - 2));
  const u32 q = STRIDED ? (t >> LOG_C) : (t & ((1u << (S - 2)) - 1));
  const u32 P0 = STRIDED ? a.lo_bit : 0;
  u32 pos0;  // in-column position of element (c, r = 0)
  bool valid = true;
  u64 col;
  if constexpr (STRIDED) {
    const u32 tpc = L - LAT_TL;  // 2^tpc tiles per column
    col = bid >> tpc;
    const u32 tc = bid & ((1u << tpc) - 1);
    const u32 nl = P0 - LOG_C;
    pos0 = ((tc >> nl) << (P0 + S)) + ((tc & ((1u << nl) - 1)) << LOG_C) + c;
  } else {
    const u32 cb = L - S;  // 2^cb chunks per column
    const u64 g = ((u64)bid << LOG_C) + c;
    valid = g < ((u64)a.num_cols << cb);
    col = valid ? g >> cb : 0;
    pos0 = (u32)(g & ((1ull << cb) - 1)) << S;
  }
  const u64 *in = col_in(a, col);
  u64 *out = col_out(a, col);
  // LDS slot of tile element (c, r)
  auto slot = [&](u32 r) {
    const u32 u = STRIDED ? r * C + c : (c << S) + r;
    return u + (u >> 4);
  };
  auto pos = [&](u32 r) { return pos0 + (r << P0); };
  const bool top = STRIDED ? TOP : (S == L);  // the pass holds the top bits: one round has no twiddle

  // first round's elements, then every round's twiddle powers (W^1..W^3 per round)
  u64 v[VALS];
  constexpr int WB0 = Sh::WB(0);
  static_for<4>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    const u64 x = valid ? gload(in + pos(window2<WB0>(q, m))) : 0;
    v[m] = a.in_canonical ? x : canon(x);
  });
  u64 tw[R][4];
  static_for<R>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    constexpr int K = Sh::K(r);
    const u32 e = round_exp(a, pos(window2<Sh::WB(r)>(q, 0)), P0 + Sh::AB(r), K);
    static_for<(1 << K)>([&](auto i_) {
      constexpr int i = decltype(i_)::value;
      if constexpr (i > 0) tw[r][i] = e ? pow_w(a.tab, e * i, INV) : 1;
    });
  });
  // fused scale: natural j0 of register 0 of the scaled round and the factor of register m
  constexpr bool PERM = N2B != !INV;  // the scaled round's layout is bit-reversed
  auto scale4 = [&](u32 p0) {
    const u32 j0 = PERM ? bitrev32(p0, L) : p0;
    const u64 f = scale_base(a, j0, INV);
    static_for<4>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      const u64 tm = a.scale_tab[4 * (PERM ? cbitrev(m, 2) : m)];
      v[m] = mul(v[m], a.scale_coset ? mul(f, tm) : tm);
    });
  };
  if constexpr (SCALE && !INV) scale4(pos(window2<WB0>(q, 0)));

  static_for<R>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    constexpr int K = Sh::K(r), WB = Sh::WB(r);
    constexpr int FIRST = 1 << (2 - K);  // registers m >= FIRST carry W^(m >> (2 - K))
    if constexpr (r > 0) {
      static_for<4>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        v[m] = lds[slot(window2<WB>(q, m))];
      });
    }
    const bool twiddle = !(top && r == Sh::TOP);
    if (N2B && twiddle) {
      static_for<4>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        if constexpr (m >= FIRST) v[m] = mul(v[m], tw[r][m >> (2 - K)]);
      });
    }
    static_for<FIRST>([&](auto lo_) {
      constexpr int lo = decltype(lo_)::value;
      if constexpr (N2B)
        dit_group<K, FIRST, lo, INV>(v);
      else
        dif_group<K, FIRST, lo, INV>(v);
    });
    if (!N2B && twiddle) {
      static_for<4>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        if constexpr (m >= FIRST) v[m] = mul(v[m], tw[r][m >> (2 - K)]);
      });
    }
    if constexpr (r + 1 < R) {
      // a round writes the slots it read, so the only barrier is before the next layout's reads
      static_for<4>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        lds[slot(window2<WB>(q, m))] = v[m];
      });
      __syncthreads();
    } else {
      if constexpr (SCALE && INV) scale4(pos(window2<WB>(q, 0)));
      if (valid) {
        static_for<4>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          gstore(out + pos(window2<WB>(q, m)), v[m]);
        });
      }
    }
  });
}

template <int S, bool N2B, bool INV, bool STRIDED, bool TOP, bool SCALE>
__global__ __launch_bounds__(THREADS) void ntt_pass_lat(PassArgs a) {
  lat_body<S, N2B, INV, STRIDED, TOP, SCALE>(a, blockIdx.x);
}
// a group of batches (batch plans), as the grouped throughput passes
template <int S, bool N2B, bool INV, bool STRIDED, bool TOP, bool SCALE>
__global__ __launch_bounds__(THREADS) void ntt_pass_lat_group(const PassArgs *jobs, const u32 *wg_start, u32 njobs) {
  const u32 j = group_of(wg_start, njobs, blockIdx.x);
  lat_body<S, N2B, INV, STRIDED, TOP, SCALE>(jobs[j], blockIdx.x - wg_start[j]);
}

}  // namespace ntt
}  // namespace gl
