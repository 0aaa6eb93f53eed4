// ntt_kernels.h -- batched Goldilocks NTT passes for gfx950.
//
// An NTT of size 2^L is the radix-2 Cooley-Tukey flow of the reference (natural -> bit-reversed
// "n2b", crates/boojum-cuda/native/ntt_n2b.cuh; bit-reversed -> natural "b2n",
// ntt_b2n.cuh) regrouped into radix-16 register blocks.  A block of K <= 4 consecutive stages
// acting on index bits [P, P + K) is
//     n2b: x_i *= W^i, then a 2^K-point DIT on the K bits,
//     b2n: a 2^K-point DIF on the K bits, then x_i *= W^i,
// with i the K-bit local index and W = omega_{2^(L-P)}^bitrev_{L-P-K}(index >> (P + K)).  The
// local 2^K-point transforms only use powers of omega_16, which are powers of two mod p
// (omega_64 = 2^39), so their twiddles are shifts.
//
// A pass covers S <= 12 index bits with one 4096-element tile per 256-thread workgroup,
// 16 elements per thread in registers, LDS transposes between radix-16 rounds:
//   * contiguous pass: bits [0, S), each tile holds 4096/2^S contiguous chunks of 2^S;
//   * strided pass:    bits [P0, P0 + S), S <= 8, each tile holds 2^S rows of 4096/2^S >= 16
//                      consecutive elements (128 B+ runs), read and written in place.
// n2b runs its strided passes (high bits first) then the contiguous pass; b2n runs the
// contiguous pass first.  The forward coset scale s^j is fused into the first pass's loads and
// the inverse n^-1 s^-j scale into the last pass's stores.
#pragma once

#include <utility>

#include "gl_tables.h"

namespace gl {
namespace ntt {

constexpr int TILE_LOG = 12;
constexpr int TILE = 1 << TILE_LOG;
constexpr int THREADS = 256;
constexpr int VALS = 16;

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int cbitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}

// omega_{2^m}^x = 2^E (mod p) for m <= 6, since omega_64 = 2^39 and 2^192 = 1.
constexpr int local_exp(int m, int x, bool inv) {
  const int v = (39 * (1 << (6 - m)) * x) % 192;
  return inv ? (192 - v) % 192 : v;
}

__device__ __forceinline__ u32 bitrev32(u32 x, u32 bits) { return bits ? (__brev(x) >> (32 - bits)) : 0; }

// (a, b) <- (a + 2^E b, a - 2^E b); 2^96 = -1 turns E >= 96 into a swapped butterfly.
template <int E>
__device__ __forceinline__ void bfly_dit(u64 &a, u64 &b) {
  const u64 a0 = a;
  if constexpr (E >= 96) {
    const u64 t = mul_pow2<E - 96>(b);
    a = sub(a0, t);
    b = add(a0, t);
  } else {
    const u64 t = mul_pow2<E>(b);
    a = add(a0, t);
    b = sub(a0, t);
  }
}

// (a, b) <- (a + b, 2^E (a - b))
template <int E>
__device__ __forceinline__ void bfly_dif(u64 &a, u64 &b) {
  const u64 s = add(a, b);
  if constexpr (E >= 96)
    b = mul_pow2<E - 96>(sub(b, a));
  else
    b = mul_pow2<E>(sub(a, b));
  a = s;
}

// 2^K-point natural -> bit-reversed DIT on v[BASE + i * STRIDE], i < 2^K
template <int K, int STRIDE, int BASE, bool INV>
__device__ __forceinline__ void dit_group(u64 (&v)[VALS]) {
  static_for<K>([&](auto t_) {
    constexpr int t = decltype(t_)::value;
    constexpr int half = 1 << (K - 1 - t);
    static_for<(1 << t)>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      constexpr int E = local_exp(t + 1, cbitrev(q, t), INV);
      static_for<half>([&](auto j_) {
        constexpr int j = q * 2 * half + decltype(j_)::value;
        bfly_dit<E>(v[BASE + j * STRIDE], v[BASE + (j + half) * STRIDE]);
      });
    });
  });
}

// 2^K-point bit-reversed -> natural DIF on v[BASE + i * STRIDE], i < 2^K
template <int K, int STRIDE, int BASE, bool INV>
__device__ __forceinline__ void dif_group(u64 (&v)[VALS]) {
  static_for<K>([&](auto t_) {
    constexpr int t = decltype(t_)::value;
    constexpr int half = 1 << t;
    static_for<(1 << (K - 1 - t))>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      constexpr int E = local_exp(K - t, cbitrev(q, K - t - 1), INV);
      static_for<half>([&](auto j_) {
        constexpr int j = q * 2 * half + decltype(j_)::value;
        bfly_dif<E>(v[BASE + j * STRIDE], v[BASE + (j + half) * STRIDE]);
      });
    });
  });
}

// One radix-2^K round over all 16 registers: register m = i * 2^(4-K) + m_lo, i active.
template <int K, bool N2B, bool INV>
__device__ __forceinline__ void local_round(u64 (&v)[VALS]) {
  static_for<(1 << (4 - K))>([&](auto lo_) {
    constexpr int lo = decltype(lo_)::value;
    if constexpr (N2B)
      dit_group<K, (1 << (4 - K)), lo, INV>(v);
    else
      dif_group<K, (1 << (4 - K)), lo, INV>(v);
  });
}

// v[m] *= W^(m >> (4-K)) for W = omega_{2^30}^(+-e)
template <int K>
__device__ __forceinline__ void apply_twiddles(u64 (&v)[VALS], u64 w) {
  constexpr int NW = 1 << K;
  u64 pw[NW];
  pw[1] = w;
  static_for<NW>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    if constexpr (i >= 2) pw[i] = (i & 1) ? mul(pw[i - 1], w) : sqr(pw[i / 2]);
  });
  static_for<VALS>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    constexpr int i = m >> (4 - K);
    if constexpr (i > 0) v[m] = mul(v[m], pw[i]);
  });
}

struct PassArgs {
  const u64 *in;
  u64 *out;
  u64 stride_in;
  u64 stride_out;
  u32 num_cols;
  u32 log_n;      // L
  u32 lo_bit;     // P0, lowest index bit this pass transforms
  u32 bits;       // S
  u32 pre_scale;  // forward coset: inputs *= s^j
  u32 post_scale; // inverse: outputs *= n^-1 (s^-j when coset)
  u32 coset;      // log_extension_degree > 0
  u32 coset_mul;  // coset_index << (30 - L - log_ext): s^j = 7^j omega_{2^30}^(j * coset_mul)
  u32 in_bitrev;  // first pass only: natural element j is read from position bitrev_L(j)
  Tables tab;
};

// s^j or n^-1 s^-j for natural index j
__device__ __forceinline__ u64 scale_factor(const PassArgs &a, u32 j, bool inverse) {
  if (!inverse) {
    u64 f = table_pow(a.tab.g, j);
    if (a.coset_mul) f = mul(f, pow_w(a.tab, j * a.coset_mul, false));
    return f;
  }
  u64 f = a.tab.inv_sizes[a.log_n];
  if (a.coset) {
    f = mul(f, table_pow(a.tab.g_inv, j));
    if (a.coset_mul) f = mul(f, pow_w(a.tab, j * a.coset_mul, true));
  }
  return f;
}

// index bits of a for register m in the round whose 4-bit window starts at bit wb
__device__ __forceinline__ u32 window_a(u32 q, u32 wb, u32 m) {
  return ((q >> wb) << (wb + 4)) | (m << wb) | (q & ((1u << wb) - 1));
}

__device__ __forceinline__ u32 lds_pad(u32 u) { return u + (u >> 4); }

// Runs round `r` of a pass (its window, its K, twiddles).  gidx0 = in-column index of the
// thread's element with window bits zero.
template <bool N2B, bool INV>
__device__ __forceinline__ void run_round(u64 (&v)[VALS], const PassArgs &a, u32 K, u32 wb, u32 ab,
                                          u32 gidx0) {
  const u32 L = a.log_n;
  const u32 P = a.lo_bit + ab;
  const u32 hi_bits = L - P - K;  // bits above the active ones
  u64 w = 1;
  bool twiddle = false;
  if (hi_bits) {
    const u32 r = gidx0 >> (P + K);
    const u32 e = bitrev32(r, hi_bits) << (OMEGA_LOG_ORDER - (L - P));
    if (e) {
      w = pow_w(a.tab, e, INV);
      twiddle = true;
    }
  }
  switch (K) {
    case 4:
      if (N2B && twiddle) apply_twiddles<4>(v, w);
      local_round<4, N2B, INV>(v);
      if (!N2B && twiddle) apply_twiddles<4>(v, w);
      break;
    case 3:
      if (N2B && twiddle) apply_twiddles<3>(v, w);
      local_round<3, N2B, INV>(v);
      if (!N2B && twiddle) apply_twiddles<3>(v, w);
      break;
    case 2:
      if (N2B && twiddle) apply_twiddles<2>(v, w);
      local_round<2, N2B, INV>(v);
      if (!N2B && twiddle) apply_twiddles<2>(v, w);
      break;
    default:
      if (N2B && twiddle) apply_twiddles<1>(v, w);
      local_round<1, N2B, INV>(v);
      if (!N2B && twiddle) apply_twiddles<1>(v, w);
      break;
  }
}

// round schedule: n2b goes from the top bits down, the partial (K < 4) round first;
// b2n goes from the bottom up, the partial round last.
__device__ __forceinline__ void round_shape(bool n2b, u32 S, u32 r, u32 &K, u32 &wb, u32 &ab) {
  const u32 k0 = S & 3;
  const u32 full = S >> 2;
  if (n2b) {
    if (k0 && r == 0) {
      K = k0;
      wb = S - 4;
      ab = S - k0;
    } else {
      const u32 idx = k0 ? r - 1 : r;  // idx-th full window from the top
      K = 4;
      wb = S - k0 - 4 * (idx + 1);
      ab = wb;
    }
  } else {
    if (r < full) {
      K = 4;
      wb = 4 * r;
      ab = wb;
    } else {
      K = k0;
      wb = S - 4;
      ab = S - k0;
    }
  }
}

template <bool N2B, bool INV>
__global__ __launch_bounds__(THREADS) void ntt_pass_strided(PassArgs a) {
  __shared__ u64 lds[TILE];
  const u32 t = threadIdx.x;
  const u32 L = a.log_n, S = a.bits, P0 = a.lo_bit;
  const u32 log_c = TILE_LOG - S;
  const u32 tiles_per_col = 1u << (L - TILE_LOG);
  const u32 col = blockIdx.x >> (L - TILE_LOG);
  const u32 tc = blockIdx.x & (tiles_per_col - 1);
  const u32 nl = P0 - log_c;
  const u32 lblk = tc & ((1u << nl) - 1);
  const u32 h = tc >> nl;
  const u32 c = t & ((1u << log_c) - 1);
  const u32 q = t >> log_c;
  const u32 base = (h << (P0 + S)) + (lblk << log_c) + c;  // in-column index of (a = 0, c)
  const u64 *in = a.in + (u64)col * a.stride_in;
  u64 *out = a.out + (u64)col * a.stride_out;

  const u32 rounds = (S >> 2) + ((S & 3) ? 1 : 0);
  u64 v[VALS];
  u32 K, wb, ab;
  round_shape(N2B, S, 0, K, wb, ab);
  if (a.in_bitrev) {
    // Top pass (P0 + S == L) over an input stored bit-reversed: the natural rows a of the
    // tile's column l sit, permuted, in the contiguous chunk [bitrev(l) 2^S, +2^S).  Stage
    // them through LDS with coalesced chunk reads.
    const u32 l0 = lblk << log_c;
    static_for<VALS>([&](auto i_) {
      constexpr int i = decltype(i_)::value;
      const u32 e = t + THREADS * i;
      const u32 cl = e >> S, off = e & ((1u << S) - 1);
      const u32 l = l0 + cl;
      const u32 arow = bitrev32(off, S);
      u64 x = canon(in[(bitrev32(l, P0) << S) + off]);
      const u32 j = (arow << P0) + l;
      if (a.pre_scale) x = mul(x, scale_factor(a, N2B ? j : bitrev32(j, L), false));
      lds[(arow << log_c) + cl] = x;
    });
    __syncthreads();
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      v[m] = lds[(window_a(q, wb, m) << log_c) + c];
    });
    __syncthreads();
  } else {
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      const u32 idx = base + (window_a(q, wb, m) << P0);
      u64 x = canon(in[idx]);
      if (a.pre_scale) x = mul(x, scale_factor(a, N2B ? idx : bitrev32(idx, L), false));
      v[m] = x;
    });
  }
  for (u32 r = 0;; r++) {
    run_round<N2B, INV>(v, a, K, wb, ab, base + (window_a(q, wb, 0) << P0));
    if (r + 1 == rounds) break;
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      lds[(window_a(q, wb, m) << log_c) + c] = v[m];
    });
    __syncthreads();
    round_shape(N2B, S, r + 1, K, wb, ab);
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      v[m] = lds[(window_a(q, wb, m) << log_c) + c];
    });
    __syncthreads();
  }
  static_for<VALS>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    const u32 idx = base + (window_a(q, wb, m) << P0);
    u64 x = v[m];
    if (a.post_scale) x = mul(x, scale_factor(a, N2B ? bitrev32(idx, L) : idx, true));
    out[idx] = x;
  });
}

template <bool N2B, bool INV>
__global__ __launch_bounds__(THREADS) void ntt_pass_contig(PassArgs a) {
  __shared__ u64 lds[TILE + TILE / 16];
  const u32 t = threadIdx.x;
  const u32 L = a.log_n, S = a.bits;
  const u32 log_c = TILE_LOG - S;
  const u32 chunk_bits = L - S;  // chunks per column = 2^chunk_bits
  const u64 total_chunks = (u64)a.num_cols << chunk_bits;
  const u64 g0 = (u64)blockIdx.x << log_c;

  // coalesced load of the tile into LDS: linear element e -> chunk e >> S, offset e & (2^S-1)
  static_for<VALS>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    const u32 e = t + THREADS * i;
    const u64 g = g0 + (e >> S);
    u64 x = 0;
    if (g < total_chunks) {
      const u64 col = g >> chunk_bits;
      const u32 idx = (u32)((g & ((1ull << chunk_bits) - 1)) << S) + (e & ((1u << S) - 1));
      x = canon(a.in[col * a.stride_in + idx]);
      // a bit-reversed input (single pass, S == L) is a permutation inside the chunk
      const u32 nat = a.in_bitrev ? bitrev32(idx, L) : idx;
      if (a.pre_scale) x = mul(x, scale_factor(a, N2B ? nat : bitrev32(nat, L), false));
      lds[lds_pad(a.in_bitrev ? ((e >> S) << S) + nat : e)] = x;
      return;  // next element (this is the static_for body)
    }
    lds[lds_pad(e)] = x;
  });
  __syncthreads();

  const u32 q = t & ((1u << (S - 4)) - 1);
  const u32 c = t >> (S - 4);
  const u64 my_chunk = g0 + c;
  const u32 hbase = (u32)((my_chunk & ((1ull << chunk_bits) - 1)) << S);
  const u32 rounds = (S >> 2) + ((S & 3) ? 1 : 0);
  u64 v[VALS];
  u32 K, wb, ab;
  round_shape(N2B, S, 0, K, wb, ab);
  for (u32 r = 0;; r++) {
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      v[m] = lds[lds_pad((c << S) + window_a(q, wb, m))];
    });
    __syncthreads();
    run_round<N2B, INV>(v, a, K, wb, ab, hbase + window_a(q, wb, 0));
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      lds[lds_pad((c << S) + window_a(q, wb, m))] = v[m];
    });
    __syncthreads();
    if (r + 1 == rounds) break;
    round_shape(N2B, S, r + 1, K, wb, ab);
  }
  static_for<VALS>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    const u32 e = t + THREADS * i;
    const u64 g = g0 + (e >> S);
    if (g < total_chunks) {
      const u64 col = g >> chunk_bits;
      const u32 idx = (u32)((g & ((1ull << chunk_bits) - 1)) << S) + (e & ((1u << S) - 1));
      u64 x = lds[lds_pad(e)];
      if (a.post_scale) x = mul(x, scale_factor(a, N2B ? bitrev32(idx, L) : idx, true));
      a.out[col * a.stride_out + idx] = x;
    }
  });
}

// log_n <= 3: one thread per column, the whole transform is one local block.
template <int K, bool N2B, bool INV>
__global__ __launch_bounds__(THREADS) void ntt_tiny(PassArgs a) {
  const u32 col = blockIdx.x * THREADS + threadIdx.x;
  if (col >= a.num_cols) return;
  const u64 *in = a.in + (u64)col * a.stride_in;
  u64 *out = a.out + (u64)col * a.stride_out;
  u64 v[VALS];
  static_for<(1 << K)>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    u64 x = canon(in[a.in_bitrev ? cbitrev(i, K) : i]);
    if (a.pre_scale) x = mul(x, scale_factor(a, N2B ? i : cbitrev(i, K), false));
    v[i] = x;
  });
  if constexpr (N2B)
    dit_group<K, 1, 0, INV>(v);
  else
    dif_group<K, 1, 0, INV>(v);
  static_for<(1 << K)>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    u64 x = v[i];
    if (a.post_scale) x = mul(x, scale_factor(a, N2B ? cbitrev(i, K) : i, true));
    out[i] = x;
  });
}

}  // namespace ntt
}  // namespace gl
