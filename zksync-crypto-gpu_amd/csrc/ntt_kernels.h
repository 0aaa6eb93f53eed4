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
// A pass covers S <= 12 index bits (S a template parameter, so every round shape, register
// index and LDS offset is a compile-time constant) with a 4096-element tile per 256-thread
// workgroup, 16 elements per thread in registers, LDS transposes between radix-16 rounds:
//   * contiguous pass: bits [0, S), each tile holds 4096/2^S contiguous chunks of 2^S;
//   * strided pass:    bits [P0, P0 + S), S <= 10, each tile holds 2^S rows of 4096/2^S >= 4
//                      consecutive elements (128 B+ runs up to S = 8, 64 and 32 B runs for the
//                      9- and 10-bit passes that let 2^21 and 2^22 run in two passes), read
//                      and written in place.
// A workgroup repeats its tile over several columns (tiles that share twiddles), computing the
// round twiddles and scale factors once and keeping them in registers.
// n2b runs its strided passes (high bits first) then the contiguous pass; b2n runs the
// contiguous pass first.  The forward coset scale s^j is fused into the first pass and the
// inverse n^-1 s^-j scale into the last pass.  The natural indices j of a thread's 16 elements
// there are j0 + pi(m) 2^(L-4) (pi the identity or a 4-bit reversal), so the factor is
// F(j0) * T[pi(m)] with T[m] = (s^(+-2^(L-4)))^m (times n^-1) built on the host per launch.
#pragma once

#include <utility>

#include "gl_phased.h"
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

// (a, b) <- (a + 2^E b, a - 2^E b); a negative shift factor (2^96 = -1, and 2^E = -2^-(96-E)
// for 64 <= E < 96) swaps the add and the subtract instead of negating.
template <int E>
__device__ __forceinline__ void bfly_dit(u64 &a, u64 &b) {
  const u64 a0 = a;
  bool negate;
  const u64 t = mul_pow2_signed<E>(b, negate);
  if (negate) {
    a = sub(a0, t);
    b = add(a0, t);
  } else {
    a = add(a0, t);
    b = sub(a0, t);
  }
}

// (a, b) <- (a + b, 2^E (a - b))
template <int E>
__device__ __forceinline__ void bfly_dif(u64 &a, u64 &b) {
  constexpr int F = E % 96;
  constexpr bool flip = (E >= 96) != (F >= 64);  // sign of the factor mul_pow2_signed returns
  const u64 s = add(a, b);
  bool negate;
  b = mul_pow2_signed<E>(flip ? sub(b, a) : sub(a, b), negate);
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

// The 8 butterflies of stage t of one radix-2^K round over all 16 registers (register
// m = i * 2^(4-K) + lo, i the K-bit local index), numbered b = lo 2^(K-1) + q half + jj: the
// registers they pair and their factor 2^E (the same flows as dit_group / dif_group).
template <int K, bool N2B, bool INV>
struct RoundStage {
  static constexpr int STRIDE = 1 << (4 - K);
  static constexpr int half(int t) { return N2B ? 1 << (K - 1 - t) : 1 << t; }
  static constexpr int j(int t, int b) {
    const int r = b & ((1 << (K - 1)) - 1);
    return (r / half(t)) * 2 * half(t) + r % half(t);
  }
  static constexpr int ia(int t, int b) { return (b >> (K - 1)) + j(t, b) * STRIDE; }
  static constexpr int ib(int t, int b) { return (b >> (K - 1)) + (j(t, b) + half(t)) * STRIDE; }
  static constexpr int E(int t, int b) {
    const int q = (b & ((1 << (K - 1)) - 1)) / half(t);
    return N2B ? local_exp(t + 1, cbitrev(q, t), INV) : local_exp(K - t, cbitrev(q, K - t - 1), INV);
  }
  static constexpr int phases(int t, int b) {
    const int e = E(t, b), f = e % 96;
    return ph::ADDSUB_PHASES + ph::shift_phases(f);
  }
};

// One radix-2^K round over all 16 registers, one butterfly after another (register
// m = i * 2^(4-K) + m_lo, i active).
template <int K, bool N2B, bool INV>
__device__ __forceinline__ void local_round_seq(u64 (&v)[VALS]) {
  static_for<(1 << (4 - K))>([&](auto lo_) {
    constexpr int lo = decltype(lo_)::value;
    if constexpr (N2B)
      dit_group<K, (1 << (4 - K)), lo, INV>(v);
    else
      dif_group<K, (1 << (4 - K)), lo, INV>(v);
  });
}

// The same round stage by stage; each stage's 8 independent butterflies run ph::ILP at a time,
// phase-interleaved.
template <int K, bool N2B, bool INV>
__device__ __forceinline__ void local_round_ilv(u64 (&v)[VALS]) {
  using St = RoundStage<K, N2B, INV>;
  static_assert(8 % ph::ILP == 0, "interleave width must divide a stage's 8 butterflies");
  static_for<K>([&](auto t_) {
    constexpr int t = decltype(t_)::value;
    static_for<8 / ph::ILP>([&](auto g_) {
      constexpr int g = decltype(g_)::value;
      ph::Bfly st[ph::ILP];
      static_for<ph::ILP>([&](auto j_) {
        constexpr int b = g * ph::ILP + decltype(j_)::value;
        st[decltype(j_)::value].a = v[St::ia(t, b)];
        st[decltype(j_)::value].b = v[St::ib(t, b)];
      });
      constexpr int MAXP = [] {
        int m = 0;
        for (int k = 0; k < ph::ILP; k++) m = St::phases(t, g * ph::ILP + k) > m ? St::phases(t, g * ph::ILP + k) : m;
        return m;
      }();
      static_for<MAXP>([&](auto p_) {
        constexpr int p = decltype(p_)::value;
        static_for<ph::ILP>([&](auto j_) {
          constexpr int b = g * ph::ILP + decltype(j_)::value;
          constexpr int e = St::E(t, b);
          if constexpr (p < St::phases(t, b)) {
            if constexpr (N2B)
              ph::dit_phase<e, p>(st[decltype(j_)::value]);
            else
              ph::dif_phase<e, p>(st[decltype(j_)::value]);
          }
        });
      });
      static_for<ph::ILP>([&](auto j_) {
        constexpr int b = g * ph::ILP + decltype(j_)::value;
        v[St::ia(t, b)] = st[decltype(j_)::value].a;
        v[St::ib(t, b)] = st[decltype(j_)::value].b;
      });
    });
  });
}

// W^0 .. W^(2^K - 1)
template <int K>
__device__ __forceinline__ void twiddle_powers(u64 (&pw)[VALS], u64 w) {
  pw[0] = 1;
  pw[1] = w;
  static_for<(1 << K)>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    if constexpr (i >= 2) pw[i] = (i & 1) ? mul(pw[i - 1], w) : sqr(pw[i / 2]);
  });
}

// v[m] *= W^(m >> (4-K)) for the registers with a factor other than W^0; ILV: interleaved
template <int K, bool ILV>
__device__ __forceinline__ void apply_twiddles(u64 (&v)[VALS], const u64 (&pw)[VALS]) {
  constexpr int FIRST = 1 << (4 - K);
  if constexpr (ILV) {
    ph::mul_batch<VALS - FIRST>([&](auto i) -> u64 & { return v[FIRST + decltype(i)::value]; },
                                [&](auto i) { return pw[(FIRST + decltype(i)::value) >> (4 - K)]; });
  } else {
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      if constexpr (m >= FIRST) v[m] = mul(v[m], pw[m >> (4 - K)]);
    });
  }
}

// v[m] *= f(m) for all 16 registers; ILV: interleaved
template <bool ILV, typename F>
__device__ __forceinline__ void mul_all(u64 (&v)[VALS], F &&f) {
  if constexpr (ILV) {
    ph::mul_batch<VALS>([&](auto i) -> u64 & { return v[decltype(i)::value]; },
                        [&](auto i) { return f(decltype(i)::value); });
  } else {
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      v[m] = mul(v[m], f(m));
    });
  }
}

template <int K, bool N2B, bool INV, bool ILV>
__device__ __forceinline__ void local_round(u64 (&v)[VALS]) {
  if constexpr (ILV)
    local_round_ilv<K, N2B, INV>(v);
  else
    local_round_seq<K, N2B, INV>(v);
}

enum : u32 { SCALE_NONE = 0, SCALE_PRE = 1, SCALE_POST = 2 };

struct PassArgs {
  const u64 *in;
  u64 *out;
  u64 stride_in;
  u64 stride_out;
  u32 num_cols;
  u32 log_n;             // L
  u32 lo_bit;            // P0, lowest index bit this pass transforms
  u32 tile_count;        // tiles in the batch
  u32 log_tile_stride;   // tiles sharing twiddles are 2^log_tile_stride apart
  u32 reps;              // tiles (columns) per workgroup
  u32 scale;             // SCALE_*: forward coset pre-scale / inverse post-scale
  u32 scale_coset;       // the factor depends on j (log_extension_degree > 0)
  u32 coset_mul;         // s^j = 7^j omega_{2^30}^(j * coset_mul)
  u32 in_bitrev;         // first n2b pass: natural element j is read from position bitrev_L(j)
  u32 in_canonical;      // the input was written by a pass of this library (values < p): loads
                         // skip the reduction that a caller's raw u64 input needs
  // column k starts at in_cols[k] / out_cols[k] when these lists are given (columns gathered
  // from separate buffers), else at in + k * stride_in / out + k * stride_out
  const u64 *const *in_cols;
  u64 *const *out_cols;
  u64 scale_tab[VALS];   // T[m]
  Tables tab;
};

__device__ __forceinline__ const u64 *col_in(const PassArgs &a, u64 col) {
  return a.in_cols ? gload(a.in_cols + col) : a.in + col * a.stride_in;
}
__device__ __forceinline__ u64 *col_out(const PassArgs &a, u64 col) {
  return a.out_cols ? gload(a.out_cols + col) : a.out + col * a.stride_out;
}

// F(j0) = s^j0 (forward) or s^-j0 (inverse), or 1 without coset
__device__ __forceinline__ u64 scale_base(const PassArgs &a, u32 j0, bool inverse) {
  if (!a.scale_coset) return 1;
  u64 f = table_pow(inverse ? a.tab.g_inv : a.tab.g, j0);
  if (a.coset_mul) f = mul(f, pow_w(a.tab, j0 * a.coset_mul, inverse));
  return f;
}

template <bool PERM>
__device__ __forceinline__ void scale_factors(u64 (&fac)[VALS], const PassArgs &a, u32 j0, bool inverse) {
  const u64 f = scale_base(a, j0, inverse);
  static_for<VALS>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    const u64 tm = a.scale_tab[PERM ? cbitrev(m, 4) : m];
    fac[m] = a.scale_coset ? mul(f, tm) : tm;
  });
}

// index bits of a for register m in the round whose 4-bit window starts at bit WB
template <int WB>
__device__ __forceinline__ u32 window_a(u32 q, int m) {
  return ((q >> WB) << (WB + 4)) | ((u32)m << WB) | (q & ((1u << WB) - 1));
}

__device__ __forceinline__ u32 lds_pad(u32 u) { return u + (u >> 4); }

// Round shapes of an S-bit pass: n2b goes from the top bits down (partial round first),
// b2n from the bottom up (partial round last).  K = stages, WB = window base, AB = lowest
// active bit (the active bits are the top K of the window).
template <int S, bool N2B>
struct Shape {
  static constexpr int K0 = S & 3;
  static constexpr int FULL = S >> 2;
  static constexpr int ROUNDS = FULL + (K0 ? 1 : 0);
  static constexpr int K(int r) {
    if (N2B) return (K0 && r == 0) ? K0 : 4;
    return r < FULL ? 4 : K0;
  }
  static constexpr int WB(int r) {
    if (N2B) return (K0 && r == 0) ? S - 4 : S - K0 - 4 * ((K0 ? r - 1 : r) + 1);
    return r < FULL ? 4 * r : S - 4;
  }
  static constexpr int AB(int r) { return WB(r) + 4 - K(r); }
  static constexpr int TOP = N2B ? 0 : ROUNDS - 1;  // the round acting on the pass's top bits
};

// W for a round whose active bits are [P, P + K) of an L-bit index; gidx = any in-column
// index of the thread's elements in that round.
__device__ __forceinline__ u32 round_exp(const PassArgs &a, u32 gidx, u32 P, u32 K) {
  const u32 L = a.log_n;
  const u32 hi_bits = L - P - K;
  if (!hi_bits) return 0;
  return bitrev32(gidx >> (P + K), hi_bits) << (OMEGA_LOG_ORDER - (L - P));
}
__device__ __forceinline__ u64 round_twiddle(const PassArgs &a, u32 gidx, u32 P, u32 K, bool inverse) {
  const u32 e = round_exp(a, gidx, P, K);
  return e ? pow_w(a.tab, e, inverse) : 1;
}

// ILV: the round's products and butterflies phase-interleaved (the strided pass, where that
// measured faster) or one after another (the contiguous pass, where it measured slower)
template <int K, bool N2B, bool INV, bool ILV>
__device__ __forceinline__ void do_round(u64 (&v)[VALS], const u64 (&pw)[VALS], bool twiddle) {
  if (N2B && twiddle) apply_twiddles<K, ILV>(v, pw);
  local_round<K, N2B, INV, ILV>(v);
  if (!N2B && twiddle) apply_twiddles<K, ILV>(v, pw);
}

// Loads/stores address a column as a uniform base plus a 32-bit per-thread byte offset
// (columns hold at most 2^28 elements), so the per-element addresses the column loop reuses are
// immediate or scalar offsets rather than registers.
__device__ __forceinline__ const u64 *at(const u64 *base, u32 byte_off) {
  return (const u64 *)((const char *)base + byte_off);
}
__device__ __forceinline__ u64 *at(u64 *base, u32 byte_off) { return (u64 *)((char *)base + byte_off); }
// a pointer every lane of the workgroup holds, moved to scalar registers
template <typename T>
__device__ __forceinline__ T *in_sgprs(T *p) {
  const u64 x = (u64)p;
  const u32 lo = __builtin_amdgcn_readfirstlane((u32)x);
  const u32 hi = __builtin_amdgcn_readfirstlane((u32)(x >> 32));
  return (T *)(((u64)hi << 32) | lo);
}

// TOP: the pass holds the index's top bits (P0 + S == L), so one round has no twiddle.
// SCALE: fused n2b pre-scale / b2n post-scale (only the top pass is first / last).
// GATHER: the n2b top pass reads a column stored bit-reversed.
// Strided variants whose next-column prefetch (16 more registers) would push them past 128 VGPRs:
// the gathered pass of 6..8 bits, plain passes of 6..8 bits below the top and of 9..10 bits at the
// top.  Their default build loads each column when its turn comes instead of while the previous
// one is transformed and runs four waves per SIMD instead of three (the 4- and 5-bit gathered
// ones, nearer the HBM bound, keep the prefetch).  Four-waves workgroups fill a CU's LDS, so a
// launch that shares the GPU with other streams' work (serve batches) takes the prefetching
// three-waves build, which leaves room for theirs.
template <int S, bool TOP, bool GATHER>
constexpr bool strided_four_waves() {
  return GATHER ? S >= 6 : ((!TOP && S >= 6 && S <= 8) || (TOP && S >= 9));
}

template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER, bool PREFETCH>
__device__ __forceinline__ void strided_body(const PassArgs &a, const u32 bid) {
  static_assert(TOP || !SCALE, "only the top pass scales");
  static_assert(!GATHER || (N2B && TOP), "only the n2b top pass gathers");
  using Sh = Shape<S, N2B>;
  constexpr int R = Sh::ROUNDS;
  constexpr int LOG_C = TILE_LOG - S;  // consecutive l values per tile (>= 4)
  static_assert(S >= 4 && S <= 10, "strided passes cover 4..10 bits");
  constexpr int C = 1 << LOG_C;
  constexpr int ROW = C + 1;           // padded LDS row (odd stride in u64)
  // the gathered pass stages row R at R ROW + (R >> GPAD) (S > 6, see below)
  constexpr int GPAD = GATHER && S > 6 ? S - 6 : -1;
  __shared__ u64 lds[(1 << S) * ROW + (GPAD >= 0 ? 64 : 0)];
  const u32 t = threadIdx.x;
  const u32 L = a.log_n, P0 = a.lo_bit;
  const u32 tc = bid & ((1u << a.log_tile_stride) - 1);
  const u32 grp = bid >> a.log_tile_stride;
  const u32 c = t & (C - 1);
  const u32 q = t >> LOG_C;
  const u32 nl = P0 - LOG_C;
  const u32 lblk = tc & ((1u << nl) - 1);
  const u32 h = tc >> nl;
  const u32 base = (h << (P0 + S)) + (lblk << LOG_C) + c;  // in-column index of (a = 0, c)
  (void)L;

  u64 tw[R][VALS];
  static_for<R>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    if constexpr (!(TOP && r == Sh::TOP)) {
      const u64 w = round_twiddle(a, base + (window_a<Sh::WB(r)>(q, 0) << P0), P0 + Sh::AB(r), Sh::K(r), INV);
      twiddle_powers<Sh::K(r)>(tw[r], w);
    }
  });
  u64 fac[VALS];
  // n2b pre-scale on the first (top) pass, b2n post-scale on the last (top) pass; both act
  // on the window [S-4, S): natural j = base + a 2^P0 with a = window_a<S-4>(q, m)
  if constexpr (SCALE) scale_factors<false>(fac, a, base + (window_a<S - 4>(q, 0) << P0), INV);

  constexpr int WB0 = Sh::WB(0);
  constexpr int WBL = Sh::WB(R - 1);
  const u32 off0 = (base + (window_a<WB0>(q, 0) << P0)) * 8u;  // byte offset of register 0
  const u32 offL = (base + (window_a<WBL>(q, 0) << P0)) * 8u;
  const u32 lds0 = window_a<WB0>(q, 0) * ROW + c;
  constexpr bool gather = GATHER;
  const u32 l0 = lblk << LOG_C;
  // the next column's elements are loaded while the current one is transformed (PREFETCH below)
  u64 nxt[VALS];
  // top pass over a column stored bit-reversed: natural row r of column l sits at
  // bitrev(l) 2^S + bitrev_S(r), so column l is one contiguous 2^S-element chunk.  Element
  // e = t + 256 k of the tile is chunk position t mod 2^S (natural row bitrev_S(t mod 2^S)) of
  // column l0 + (t >> S) + k 2^(8-S); those three terms have disjoint bits (l0 a multiple of
  // 2^(12-S)), so bitrev_P0 of the column splits the same way, and the address is the column
  // base plus a per-thread byte offset that no column changes plus a uniform one per k.  A
  // wave reads consecutive addresses; the rows go to LDS in natural order.  For S > 6 a wave's
  // 64 rows are 2^(S-6) apart, which with the odd row stride alone would put them in 8 (S = 8)
  // or 16 (S = 7) of the 32 eight-byte LDS bank pairs: row R is staged at R ROW + (R >> (S-6)),
  // which spreads them over all 32.  The first round reads rows (m << (S-4)) | q, whose extra
  // term (q >> (S-6)) + 4 m stays linear in m.
  // Loads and stores take the column base in SGPRs and a 32-bit lane offset formed from one
  // register per column: the empty asm makes that register a new value in each iteration, so
  // LLVM neither hoists 16 zero-extended 64-bit offsets out of the column loop (which costs a
  // 64-bit add per access instead of the saddr form) nor keeps 16 64-bit bases in SGPRs (which
  // spill).
  const u32 gather_off =
      gather ? ((bitrev32(t >> S, P0) << S) | (t & ((1u << S) - 1))) * 8u : 0u;
  auto fetch = [&](u32 col) {
    const u64 *in = in_sgprs(col_in(a, col));
    if constexpr (gather) {
      static_assert(S <= 8, "the gathered pass reads 2^(8-S) columns per thread step");
      u32 o = gather_off + (bitrev32(l0, P0) << (S + 3));
      asm volatile("" : "+v"(o));
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        nxt[k] = gload(at(in, o + (bitrev32((u32)k << (8 - S), P0) << (S + 3))));
      });
    } else {
      u32 o = off0;
      asm volatile("" : "+v"(o));
      static_for<VALS>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        nxt[m] = gload(at(in, o + ((u32)m << (WB0 + P0 + 3))));
      });
    }
  };
  u32 col = grp * a.reps;
  const u32 col_end = min(col + a.reps, a.num_cols);
  if (PREFETCH && col < col_end) fetch(col);
  for (; col < col_end; col++) {
    u64 *out = in_sgprs(col_out(a, col));
    if (!PREFETCH) fetch(col);
    u64 v[VALS];
    if constexpr (gather) {
      const u32 row = bitrev32(t & ((1u << S) - 1), S);
      const u32 st0 = row * ROW + (GPAD >= 0 ? row >> GPAD : 0) + (t >> S);
      auto stage = [&](auto reduce_) {
        static_for<VALS>([&](auto k_) {
          constexpr int k = decltype(k_)::value;
          lds[st0 + (k << (8 - S))] = decltype(reduce_)::value ? canon(nxt[k]) : nxt[k];
        });
      };
      if (a.in_canonical)
        stage(std::false_type{});
      else
        stage(std::true_type{});
      __syncthreads();
      static_assert(GPAD < 0 || WB0 == GPAD + 2, "the staged rows' pad is linear in m");
      const u32 ld0 = lds0 + (GPAD >= 0 ? q >> GPAD : 0);
      static_for<VALS>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        v[m] = lds[ld0 + m * ((ROW << WB0) + (GPAD >= 0 ? 4 : 0))];
      });
      __syncthreads();
    } else {
      if (a.in_canonical) {
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = nxt[m];
        });
      } else {
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = canon(nxt[m]);
        });
      }
    }
    if (PREFETCH && col + 1 < col_end) fetch(col + 1);
    if constexpr (N2B && SCALE) mul_all<true>(v, [&](int m) { return fac[m]; });
    static_for<R>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      if constexpr (r > 0) {
        constexpr int WBP = Sh::WB(r - 1), WBR = Sh::WB(r);
        const u32 lp = window_a<WBP>(q, 0) * ROW + c;
        const u32 lr = window_a<WBR>(q, 0) * ROW + c;
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          lds[lp + m * (ROW << WBP)] = v[m];
        });
        __syncthreads();
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = lds[lr + m * (ROW << WBR)];
        });
        __syncthreads();
      }
      do_round<Sh::K(r), N2B, INV, true>(v, tw[r], !(TOP && r == Sh::TOP));
    });
    if constexpr (!N2B && SCALE) mul_all<true>(v, [&](int m) { return fac[m]; });
    u32 o = offL;
    asm volatile("" : "+v"(o));
    static_for<VALS>([&](auto m_) {
      constexpr int m = decltype(m_)::value;
      gstore(at(out, o + ((u32)m << (WBL + P0 + 3))), v[m]);
    });
  }
}

// SC: fused forward pre-scale (INV = false) or inverse post-scale (INV = true): 0 none,
// 1 one factor for every element (n^-1 of an inverse without coset), 2 per-element coset factors.
// WHOLE: columns of >= 4096 elements (a tile is one contiguous block of one column); otherwise
// a tile holds 4096 / 2^S whole columns of 2^S elements.
// PREFETCH: the next tile is loaded into registers while the current one is transformed (the
// three-waves-per-SIMD variant); otherwise each tile is loaded when its transform starts (the
// four-waves variant, which fits 128 registers).
template <int S, bool N2B, bool INV, int SC, bool WHOLE, bool PREFETCH>
__device__ __forceinline__ void contig_body(const PassArgs &a, const u32 bid) {
  constexpr bool SCALE = SC == 2;
  const u64 uniform = a.scale_tab[0];
  using Sh = Shape<S, N2B>;
  constexpr int R = Sh::ROUNDS;
  constexpr int LOG_C = TILE_LOG - S;  // chunks per tile
  __shared__ u64 lds[TILE + TILE / 16];
  const u32 t = threadIdx.x;
  const u32 L = a.log_n;
  const u32 chunk_bits = L - S;  // chunks per column = 2^chunk_bits
  const u64 total_chunks = (u64)a.num_cols << chunk_bits;
  const u32 tc = bid & ((1u << a.log_tile_stride) - 1);
  const u32 grp = bid >> a.log_tile_stride;
  const u32 q = t & ((1u << (S - 4)) - 1);
  const u32 c = t >> (S - 4);
  // chunks of the tiles this workgroup repeats over share their in-column position, hence
  // their twiddles; hbase = in-column index of the thread's chunk start
  const u64 tile0 = ((u64)grp * a.reps << a.log_tile_stride) + tc;
  const u32 hbase = (u32)((((tile0 << LOG_C) + c) & ((1ull << chunk_bits) - 1)) << S);
  // a tile of a column of >= 4096 elements is one contiguous 32 KB block of that column
  constexpr bool whole = WHOLE;
  const u32 tile_in_col = (u32)((tile0 << (LOG_C + S)) & ((1ull << L) - 1));  // first element

  // Twiddles.  A round on the window starting at bit WB >= 4 has at most 2^(8-WB) <= 16
  // distinct W per tile (the tile's 2^(12-S) chunks times the 2^(S-4-WB) values of q >> WB),
  // so the workgroup keeps their powers in LDS, one entry per thread to fill; a round on a
  // lower window (every thread its own W) keeps its powers in registers, across repeats.
  __shared__ u64 ltw[R][256];
  u64 tw[R][VALS];
  // an n2b inverse with one factor for every element (n^-1) folds it into the last round's
  // register twiddles (all 16, W^0 included) instead of a multiply after the round
  constexpr bool FOLD = N2B && INV && SC == 1 && R >= 2 && Sh::WB(R - 1) < 4;
  static_for<R>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    constexpr int WB = Sh::WB(r), K = Sh::K(r);
    if constexpr (WB >= 4) {
      constexpr int QB = S - 4 - WB;  // bits of q above the window
      const u32 idx = t >> K, i = t & ((1u << K) - 1);
      if (idx < (1u << (8 - WB))) {
        const u32 cc = idx >> QB, qq = (idx & ((1u << QB) - 1)) << WB;
        const u32 hb = (u32)((((tile0 << LOG_C) + cc) & ((1ull << chunk_bits) - 1)) << S);
        const u32 e = round_exp(a, hb + window_a<WB>(qq, 0), Sh::AB(r), K);
        ltw[r][t] = e ? pow_w(a.tab, e * i, INV) : 1;  // W^i: omega has order 2^30 | 2^32
      }
    } else {
      const u64 w = round_twiddle(a, hbase + window_a<WB>(q, 0), Sh::AB(r), K, INV);
      twiddle_powers<K>(tw[r], w);
      if constexpr (FOLD && r == R - 1) {
        static_for<(1 << K)>([&](auto i_) {
          constexpr int i = decltype(i_)::value;
          tw[r][i] = i ? mul(tw[r][i], uniform) : uniform;
        });
      }
    }
  });
  const bool top = S == L;  // single pass: the top round has no twiddle
  // Per-element coset factors F(j0) T[pi(m)].  n2b pre (single pass) and b2n post (single pass)
  // act on the window [S-4, S), identity order; n2b post and b2n pre act on the window [0, 4),
  // natural j = bitrev_L(index).  Over whole 4096-element tiles only F stays in registers and
  // each factor is formed where it is applied (one more product per element), so the pass
  // keeps the unscaled pass's three waves per SIMD; tiles of short columns keep all 16 factors
  // (two waves per SIMD), which spills less there.
  constexpr bool PERM = N2B != !INV;
  constexpr bool AT_USE = WHOLE;
  u64 fbase = 1;
  u64 fac[VALS];
  if constexpr (SCALE) {
    const u32 j0 = PERM ? bitrev32(hbase + window_a<0>(q, 0), L) : hbase + window_a<S - 4>(q, 0);
    if constexpr (AT_USE)
      fbase = scale_base(a, j0, INV);
    else
      scale_factors<PERM>(fac, a, j0, INV);
  }
  // v[m] *= the scale factor of register m (F(j0) T[pi(m)], formed here when AT_USE), or the
  // one factor for all elements
  auto scale_all = [&](u64(&v)[VALS]) {
    if constexpr (!SCALE) {
      mul_all<false>(v, [&](int) { return uniform; });
    } else if constexpr (AT_USE) {
      mul_all<false>(v, [&](int m) { return mul(fbase, a.scale_tab[PERM ? cbitrev(m, 4) : m]); });
    } else {
      mul_all<false>(v, [&](int m) { return fac[m]; });
    }
  };

  // the next tile's elements are loaded while the current one is transformed
  u64 nxt[VALS];
  auto fetch = [&](u64 tile) {
    if constexpr (whole) {
      const u64 col = tile >> a.log_tile_stride;
      const u64 *in = col_in(a, col) + tile_in_col;
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        nxt[k] = gload(at(in + THREADS * k, t * 8u));
      });
    } else {
      const u64 g0 = tile << LOG_C;
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        const u32 e = t + THREADS * k;
        const u64 g = g0 + (e >> S);
        nxt[k] = g < total_chunks ? gload(col_in(a, g) + (e & ((1u << S) - 1))) : 0;
      });
    }
  };
  const u32 reps =
      tile0 < a.tile_count
          ? (u32)min((u64)a.reps, (a.tile_count - tile0 + (1ull << a.log_tile_stride) - 1) >> a.log_tile_stride)
          : 0;
  // An n2b pass over whole 4096-element tiles starts with the window of the top 4 bits, whose 16
  // elements per thread (t + 256 m) are the ones the thread loaded: they go straight into
  // registers, with no LDS staging and its barrier (unless the load is a bit-reversed gather).
  // The twiddle table's fill then needs a barrier of its own before the first round reads it.
  const bool direct0 = N2B && S == TILE_LOG && WHOLE && !a.in_bitrev;
  if (direct0) __syncthreads();
  if (PREFETCH && reps) fetch(tile0);
  for (u32 i = 0; i < reps; i++) {
    const u64 tile = tile0 + ((u64)i << a.log_tile_stride);
    const u64 g0 = tile << LOG_C;
    if constexpr (!PREFETCH) fetch(tile);
    // keeps the 16 factors from being hoisted out of the loop (into 32 live registers)
    if constexpr (SCALE && AT_USE) asm volatile("" : "+v"(fbase));
    u64 v[VALS];
    auto stage = [&](auto reduce_) {
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        const u32 e = t + THREADS * k;
        // lds_pad(t + 256 k) = lds_pad(t) + 272 k; a bit-reversed input (single pass) is a
        // permutation inside each column
        u32 slot = lds_pad(t) + (THREADS + THREADS / 16) * k;
        if (N2B && a.in_bitrev) slot = lds_pad(((e >> S) << S) + bitrev32(e & ((1u << S) - 1), L));
        lds[slot] = decltype(reduce_)::value ? canon(nxt[k]) : nxt[k];
      });
    };
    if (direct0) {
      if (a.in_canonical) {
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = nxt[m];
        });
      } else {
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = canon(nxt[m]);
        });
      }
    } else if (a.in_canonical) {
      stage(std::false_type{});
    } else {
      stage(std::true_type{});
    }
    if (PREFETCH && i + 1 < reps) fetch(tile + (1ull << a.log_tile_stride));
    if (!direct0) __syncthreads();
    static_for<R>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      constexpr int WB = Sh::WB(r);
      const u32 u0 = (c << S) + window_a<WB>(q, 0);
      const u32 l0 = lds_pad(u0);
      // for WB >= 4 the padding advances by 2^(WB-4) per step of m, for WB = 0 (u0 a multiple
      // of 16) not at all; partial windows at WB = 1..3 (S = 5..7) are padded per element
      constexpr bool LINEAR = WB == 0 || WB >= 4;
      constexpr u32 STEP = (1u << WB) + (WB >= 4 ? (1u << (WB >= 4 ? WB - 4 : 0)) : 0);
      auto slot = [&](int m) { return LINEAR ? l0 + m * STEP : lds_pad(u0 + ((u32)m << WB)); };
      if (r != 0 || !direct0) {
        static_for<VALS>([&](auto m_) {
          constexpr int m = decltype(m_)::value;
          v[m] = lds[slot(m)];
        });
        __syncthreads();
      }
      if (SC != 0 && !INV && r == 0) scale_all(v);
      if constexpr (WB >= 4) {
        constexpr int K = Sh::K(r), QB = S - 4 - WB;
        const u32 base = ((c << QB) | (q >> WB)) << K;
        u64 pw[VALS];
        static_for<(1 << K)>([&](auto i_) {
          constexpr int i = decltype(i_)::value;
          pw[i] = ltw[r][base + i];
        });
        do_round<K, N2B, INV, false>(v, pw, !(top && r == Sh::TOP));
      } else if constexpr (FOLD && r == R - 1) {
        constexpr int K = Sh::K(r);
        mul_all<false>(v, [&](int m) { return tw[r][m >> (4 - K)]; });
        local_round<K, N2B, INV, false>(v);
      } else {
        do_round<Sh::K(r), N2B, INV, false>(v, tw[r], !(top && r == Sh::TOP));
      }
      if (SC != 0 && INV && r == R - 1 && !FOLD) scale_all(v);
      static_for<VALS>([&](auto m_) {
        constexpr int m = decltype(m_)::value;
        lds[slot(m)] = v[m];
      });
      __syncthreads();
    });
    if constexpr (whole) {
      const u64 col = tile >> a.log_tile_stride;
      u64 *out = col_out(a, col) + tile_in_col;
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        gstore(at(out + THREADS * k, t * 8u), lds[lds_pad(t) + (THREADS + THREADS / 16) * k]);
      });
    } else {
      static_for<VALS>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        const u32 e = t + THREADS * k;
        const u64 g = g0 + (e >> S);
        if (g < total_chunks) gstore(col_out(a, g) + (e & ((1u << S) - 1)), lds[lds_pad(e)]);
      });
    }
    __syncthreads();
  }
}

// Scale factor for natural index j, computed directly (tiny transforms only).
__device__ __forceinline__ u64 tiny_factor(const PassArgs &a, u32 j, bool inverse) {
  u64 f = a.scale_tab[0];  // 1 (forward) or n^-1 (inverse)
  if (a.scale_coset) f = mul(f, scale_base(a, j, inverse));
  return f;
}

// log_n <= 3: one thread per column, the whole transform is one local block.
template <int K, bool N2B, bool INV>
__device__ __forceinline__ void tiny_body(const PassArgs &a, const u32 bid) {
  const u32 col = bid * THREADS + threadIdx.x;
  if (col >= a.num_cols) return;
  const u64 *in = col_in(a, col);
  u64 *out = col_out(a, col);
  u64 v[VALS];
  static_for<(1 << K)>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    u64 x = canon(gload(in + (a.in_bitrev ? cbitrev(i, K) : i)));
    if (a.scale == SCALE_PRE) x = mul(x, tiny_factor(a, N2B ? i : cbitrev(i, K), false));
    v[i] = x;
  });
  if constexpr (N2B)
    dit_group<K, 1, 0, INV>(v);
  else
    dif_group<K, 1, 0, INV>(v);
  static_for<(1 << K)>([&](auto i_) {
    constexpr int i = decltype(i_)::value;
    u64 x = v[i];
    if (a.scale == SCALE_POST) x = mul(x, tiny_factor(a, N2B ? cbitrev(i, K) : i, true));
    gstore(out + i, x);
  });
}

// Kernels: a single batch (PassArgs by value) or a group of batches sharing one pass shape, where
// workgroup b runs batch j with wg_start[j] <= b < wg_start[j + 1] (its block index rebased).
__device__ __forceinline__ u32 group_of(const u32 *wg_start, u32 njobs, u32 b) {
  u32 lo = 0, hi = njobs;  // wg_start is ascending; find the last entry <= b
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (wg_start[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Strided passes of at most this many bits take the XCD-contiguous tile order (measured: the
// 4-bit pass over 2^22-element columns 12-15 % faster; 6- and 8-bit passes neutral or slower),
// and so do the 9- and 10-bit passes, whose rows are runs of 8 and 4 elements (64 and 32 B): the
// neighbouring tiles that read the rest of each 128 B line then run on the same XCD, whose L2
// holds it.
#ifndef BA_XCD_REMAP_MAX_S
#define BA_XCD_REMAP_MAX_S 4
#endif
template <int S>
__device__ __forceinline__ u32 strided_tile() {
  if constexpr (S <= BA_XCD_REMAP_MAX_S || S >= 9) {
    // Workgroups are dealt round-robin over the 8 XCDs, so neighbouring tiles (neighbouring runs
    // of the same 2^S rows, 2^(log_n - S) elements apart) would run at once on all eight and
    // crowd the same HBM channels.  XCD x instead takes the contiguous tile range starting at
    // x q + min(x, r), q or q + 1 tiles (q = G / 8, r = G % 8): a bijection on [0, G).
    const u32 G = gridDim.x, b = blockIdx.x, x = b & 7u, q = G >> 3, r = G & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
  } else {
    return blockIdx.x;
  }
}
template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER>
__global__ __launch_bounds__(THREADS) void ntt_pass_strided(PassArgs a) {
  strided_body<S, N2B, INV, TOP, SCALE, GATHER, !strided_four_waves<S, TOP, GATHER>()>(a, strided_tile<S>());
}
template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER>
__global__ __launch_bounds__(THREADS) void ntt_pass_strided_group(const PassArgs *jobs, const u32 *wg_start,
                                                                  u32 njobs) {
  const u32 j = group_of(wg_start, njobs, blockIdx.x);
  strided_body<S, N2B, INV, TOP, SCALE, GATHER, !strided_four_waves<S, TOP, GATHER>()>(jobs[j],
                                                                                       blockIdx.x - wg_start[j]);
}
// the prefetching three-waves build of a four-waves variant, for launches that share the GPU
template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER>
__global__ __launch_bounds__(THREADS) void ntt_pass_strided_shared(PassArgs a) {
  strided_body<S, N2B, INV, TOP, SCALE, GATHER, true>(a, strided_tile<S>());
}
template <int S, bool N2B, bool INV, bool TOP, bool SCALE, bool GATHER>
__global__ __launch_bounds__(THREADS) void ntt_pass_strided_shared_group(const PassArgs *jobs,
                                                                         const u32 *wg_start, u32 njobs) {
  const u32 j = group_of(wg_start, njobs, blockIdx.x);
  strided_body<S, N2B, INV, TOP, SCALE, GATHER, true>(jobs[j], blockIdx.x - wg_start[j]);
}
// W: waves per SIMD the contiguous pass is compiled for.  4: no prefetch, 128 registers (fastest
// per launch over many tiles); 3: with the next-tile prefetch (fewer instructions per tile, better
// when few workgroups run, and it leaves LDS room for a concurrent launch).  The per-element-
// scaled variants over short columns keep 16 more values per thread and run at 2 either way.
// The autotuner picks W per size class (ba_tuned_plan.contig_waves).
template <int S, bool N2B, bool INV, int SC, bool WHOLE, int W>
__global__ __launch_bounds__(THREADS, SC == 2 && !WHOLE ? 2 : W) void ntt_pass_contig(PassArgs a) {
  contig_body<S, N2B, INV, SC, WHOLE, W == 3>(a, blockIdx.x);
}
template <int S, bool N2B, bool INV, int SC, bool WHOLE, int W>
__global__ __launch_bounds__(THREADS, SC == 2 && !WHOLE ? 2 : W) void ntt_pass_contig_group(const PassArgs *jobs,
                                                                                      const u32 *wg_start,
                                                                                      u32 njobs) {
  const u32 j = group_of(wg_start, njobs, blockIdx.x);
  contig_body<S, N2B, INV, SC, WHOLE, W == 3>(jobs[j], blockIdx.x - wg_start[j]);
}
template <int K, bool N2B, bool INV>
__global__ __launch_bounds__(THREADS) void ntt_tiny(PassArgs a) {
  tiny_body<K, N2B, INV>(a, blockIdx.x);
}
template <int K, bool N2B, bool INV>
__global__ __launch_bounds__(THREADS) void ntt_tiny_group(const PassArgs *jobs, const u32 *wg_start, u32 njobs) {
  const u32 j = group_of(wg_start, njobs, blockIdx.x);
  tiny_body<K, N2B, INV>(jobs[j], blockIdx.x - wg_start[j]);
}

}  // namespace ntt
}  // namespace gl
