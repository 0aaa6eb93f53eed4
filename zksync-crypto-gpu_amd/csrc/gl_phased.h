// gl_phased.h -- the NTT passes' field arithmetic in phases, for interleaving independent
// operations.
//
// A product, a DIT butterfly and a DIF butterfly are each split into a fixed sequence of
// phases (one to three single-instruction steps each, from gl_field.h's isa namespace).  A
// batch of independent operations runs phase 0 of every member, then phase 1 of every member,
// and so on, so the instructions of different operations alternate: the wait states a
// VALU-written carry needs before it is read, and the latency of each dependent step, are
// filled with other operations' work instead of idle issue slots.  The arithmetic, and so every
// result, is exactly gl::mul / gl::add / gl::sub / gl::mul_pow2_signed's.
#pragma once

#include <utility>

#include "gl_field.h"

namespace gl {
namespace ph {

// Operations interleaved at once (the rest of a batch follows in further groups): enough to
// fill the carry wait states, few enough that their temporaries fit the registers.
#ifndef BA_ILP
#define BA_ILP 4
#endif
constexpr int ILP = BA_ILP;

template <typename F, int... I>
__device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F &&f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------- product (canonical result)
struct Mul {
  u64 a, b, r;
  u64 z, y, w, R, cy, k1, k2, c, b1, bb, msub, madd, k;
  u32 r1, r2, r3, d0, d1, e0, e1, f0;
};
constexpr int MUL_PHASES = 13;

// phase P of gl::mul(s.a, s.b) -> s.r
template <int P>
__device__ __forceinline__ void mul_phase(Mul &s) {
  using namespace isa;
  if constexpr (P == 0) {
    s.z = mul32(lo(s.a), lo(s.b));
    s.y = mul32(hi(s.a), lo(s.b));
  } else if constexpr (P == 1) {
    s.y = mad(lo(s.a), hi(s.b), s.y, s.cy);
    s.w = mul32(hi(s.a), hi(s.b));
  } else if constexpr (P == 2) {
    s.r1 = add_co(hi(s.z), lo(s.y), s.k1);
  } else if constexpr (P == 3) {
    s.r2 = addc(lo(s.w), hi(s.y), s.k1, s.k2);
  } else if constexpr (P == 4) {
    s.r3 = addc0(hi(s.w), s.k2);
  } else if constexpr (P == 5) {
    s.R = mad_eps(s.r2, pack(lo(s.z), s.r1), s.c);
  } else if constexpr (P == 6) {
    s.d0 = subb(lo(s.R), s.r3, s.cy, s.b1);
  } else if constexpr (P == 7) {
    s.d1 = subb0(hi(s.R), s.b1, s.bb);
  } else if constexpr (P == 8) {
    const u64 ge = ge_p(pack(s.d0, s.d1));
    s.msub = s.bb & ~s.c;
    s.madd = (s.c & ~s.bb) | (ge & ~(s.c ^ s.bb));
  } else if constexpr (P == 9) {
    s.e0 = addc0(s.d0, s.msub, s.k);
  } else if constexpr (P == 10) {
    s.e1 = subb0(s.d1, s.msub & ~s.k);
  } else if constexpr (P == 11) {
    s.f0 = subb0(s.e0, s.madd, s.k);
  } else if constexpr (P == 12) {
    s.r = pack(s.f0, addc0(s.e1, s.madd & ~s.k));
  }
}

// For i < N (compile-time): x(i) *= y(i), where x(ic) returns a reference to the operand and
// y(ic) its multiplier (ic: std::integral_constant<int, i>), interleaved ILP at a time.
template <int N, typename X, typename Y>
__device__ __forceinline__ void mul_batch(X &&x, Y &&y) {
  constexpr int GROUPS = (N + ILP - 1) / ILP;
  sfor<GROUPS>([&](auto g_) {
    constexpr int g = decltype(g_)::value;
    constexpr int M = (g + 1) * ILP <= N ? ILP : N - g * ILP;
    Mul st[M];
    sfor<M>([&](auto j_) {
      constexpr int j = decltype(j_)::value;
      st[j].a = x(std::integral_constant<int, g * ILP + j>{});
      st[j].b = y(std::integral_constant<int, g * ILP + j>{});
    });
    sfor<MUL_PHASES>([&](auto p_) {
      sfor<M>([&](auto j_) { mul_phase<decltype(p_)::value>(st[decltype(j_)::value]); });
    });
    sfor<M>([&](auto j_) {
      constexpr int j = decltype(j_)::value;
      x(std::integral_constant<int, g * ILP + j>{}) = st[j].r;
    });
  });
}

// ---------------------------------------------------------------- shift by 2^F, canonical result
// The three forms of gl::shl / gl::shr (x any u64):
//   F in (0, 32]:  x << F = lo + hl 2^64 -> reduce2                          5 phases
//   F in (32, 64): x << F = lo + (hl + hh 2^32) 2^64 -> reduce3              9 phases
//   F in [64, 96): 2^F = -2^-(96-F): x 2^-M = q - r 2^(32-M) EPS, M = 96 - F  6 phases
constexpr int shift_phases(int F) { return F == 0 ? 0 : F <= 32 ? 5 : F < 64 ? 9 : 6; }

struct Shift {
  u64 x, t;  // in, out
  u64 lo, R, c, b1, bb, m, k, msub, madd, q, u, bw;
  u32 hl, hh, d0, d1, e0, e1, t0, r;
};

template <int F, int P>
__device__ __forceinline__ void shift_phase(Shift &s) {
  using namespace isa;
  if constexpr (F == 0) {
    if constexpr (P == 0) s.t = s.x;  // (a zero-length form: never called with P >= 0 in a batch)
  } else if constexpr (F <= 32) {
    if constexpr (P == 0) {
      s.lo = s.x << F;
      s.hl = (u32)(s.x >> (64 - F));
    } else if constexpr (P == 1) {
      s.R = mad_eps(s.hl, s.lo, s.c);
    } else if constexpr (P == 2) {
      s.m = s.c | ge_p(s.R);
    } else if constexpr (P == 3) {
      s.t0 = subb0(lo(s.R), s.m, s.k);
    } else if constexpr (P == 4) {
      s.t = pack(s.t0, addc0(hi(s.R), s.m & ~s.k));
    }
  } else if constexpr (F < 64) {
    if constexpr (P == 0) {
      s.lo = s.x << F;
      const u64 h = s.x >> (64 - F);
      s.hl = (u32)h;
      s.hh = (u32)(h >> 32);
    } else if constexpr (P == 1) {
      s.R = mad_eps(s.hl, s.lo, s.c);
    } else if constexpr (P == 2) {
      s.d0 = sub_co(lo(s.R), s.hh, s.b1);
    } else if constexpr (P == 3) {
      s.d1 = subb0(hi(s.R), s.b1, s.bb);
    } else if constexpr (P == 4) {
      const u64 ge = ge_p(pack(s.d0, s.d1));
      s.msub = s.bb & ~s.c;
      s.madd = (s.c & ~s.bb) | (ge & ~(s.c ^ s.bb));
    } else if constexpr (P == 5) {
      s.e0 = addc0(s.d0, s.msub, s.k);
    } else if constexpr (P == 6) {
      s.e1 = subb0(s.d1, s.msub & ~s.k);
    } else if constexpr (P == 7) {
      s.t0 = subb0(s.e0, s.madd, s.k);
    } else if constexpr (P == 8) {
      s.t = pack(s.t0, addc0(s.e1, s.madd & ~s.k));
    }
  } else {
    constexpr int M = 96 - F;
    if constexpr (P == 0) {
      s.q = s.x >> M;
      s.r = M == 32 ? (u32)s.x : (u32)s.x << (32 - M);
    } else if constexpr (P == 1) {
      s.u = mad_eps(s.r, 0, s.k);
    } else if constexpr (P == 2) {
      s.d0 = sub_co(lo(s.q), lo(s.u), s.c);
    } else if constexpr (P == 3) {
      s.d1 = subb(hi(s.q), hi(s.u), s.c, s.bw);
    } else if constexpr (P == 4) {
      s.t0 = addc0(s.d0, s.bw, s.k);
    } else if constexpr (P == 5) {
      s.t = pack(s.t0, subb0(s.d1, s.bw & ~s.k));
    }
  }
}

// ---------------------------------------------------------------- add and subtract
// (s, d) = (a + b, a - b), canonical, from canonical a, b: 5 phases.
struct AddSub {
  u64 a, b, s, d;
  u64 cs, c2, cd, bw, m, k, kd;
  u32 s0, s1, d0, d1;
};
constexpr int ADDSUB_PHASES = 5;

template <int P>
__device__ __forceinline__ void addsub_phase(AddSub &s) {
  using namespace isa;
  if constexpr (P == 0) {
    s.s0 = add_co(lo(s.a), lo(s.b), s.cs);
    s.d0 = sub_co(lo(s.a), lo(s.b), s.cd);
  } else if constexpr (P == 1) {
    s.s1 = addc(hi(s.a), hi(s.b), s.cs, s.c2);
    s.d1 = subb(hi(s.a), hi(s.b), s.cd, s.bw);
  } else if constexpr (P == 2) {
    s.m = s.c2 | ge_p(pack(s.s0, s.s1));
  } else if constexpr (P == 3) {
    s.s0 = subb0(s.s0, s.m, s.k);
    s.d0 = addc0(s.d0, s.bw, s.kd);
  } else if constexpr (P == 4) {
    s.s = pack(s.s0, addc0(s.s1, s.m & ~s.k));
    s.d = pack(s.d0, subb0(s.d1, s.bw & ~s.kd));
  }
}

// ---------------------------------------------------------------- butterflies
// DIT: (a, b) <- (a + 2^E b, a - 2^E b): the shift, then add/sub; a negative factor (2^96 = -1,
// 2^F = -2^-(96-F) for 64 <= F < 96) swaps the sum and the difference.
// DIF: (a, b) <- (a + b, 2^E (a - b)), the sign folded into the order of the subtraction.
template <int E>
struct PowForm {
  static constexpr int F = E % 96;
  static constexpr bool NEG = F == 0 ? E >= 96 : (F < 64 ? E >= 96 : E < 96);
  static constexpr int SP = shift_phases(F);
};

// a, b: the butterfly's operands in, its results out
struct Bfly {
  u64 a, b;
  Shift sh;
  AddSub as;
};

template <int E, int P>
__device__ __forceinline__ void dit_phase(Bfly &s) {
  using Fm = PowForm<E>;
  if constexpr (P == 0) s.sh.x = s.b, s.as.a = s.a;
  if constexpr (P < Fm::SP) shift_phase<Fm::F, P>(s.sh);
  if constexpr (P == Fm::SP) s.as.b = Fm::SP ? s.sh.t : s.sh.x;
  if constexpr (P >= Fm::SP && P < Fm::SP + ADDSUB_PHASES) addsub_phase<P - Fm::SP>(s.as);
  if constexpr (P == Fm::SP + ADDSUB_PHASES - 1) {
    s.a = Fm::NEG ? s.as.d : s.as.s;
    s.b = Fm::NEG ? s.as.s : s.as.d;
  }
}
template <int E>
constexpr int dit_phases() {
  return PowForm<E>::SP + ADDSUB_PHASES;
}

template <int E, int P>
__device__ __forceinline__ void dif_phase(Bfly &s) {
  using Fm = PowForm<E>;
  if constexpr (P == 0) {
    // 2^E (a - b) = +-2^F' (a - b): the sign the shift form carries is folded into the order
    s.as.a = Fm::NEG ? s.b : s.a;
    s.as.b = Fm::NEG ? s.a : s.b;
  }
  if constexpr (P < ADDSUB_PHASES) addsub_phase<P>(s.as);
  if constexpr (P == ADDSUB_PHASES - 1) {
    s.a = s.as.s;  // a + b either way
    s.sh.x = s.as.d;
    if constexpr (Fm::SP == 0) s.b = s.as.d;
  }
  if constexpr (P >= ADDSUB_PHASES && P < ADDSUB_PHASES + Fm::SP) shift_phase<Fm::F, P - ADDSUB_PHASES>(s.sh);
  if constexpr (Fm::SP > 0 && P == ADDSUB_PHASES + Fm::SP - 1) s.b = s.sh.t;
}
template <int E>
constexpr int dif_phases() {
  return ADDSUB_PHASES + PowForm<E>::SP;
}

}  // namespace ph
}  // namespace gl
