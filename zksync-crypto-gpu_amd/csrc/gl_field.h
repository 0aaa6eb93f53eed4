// gl_field.h -- Goldilocks field arithmetic for gfx950 (p = 2^64 - 2^32 + 1).
//
// Written from the field's special form, not from the reference's carry-chain code:
//   2^64 = 2^32 - 1 (mod p),   2^96 = -1 (mod p),   2^192 = 1 (mod p).
// Every function here takes canonical operands (< p) unless it says otherwise and returns a
// canonical result, so stored results compare bit-exactly with the CPU oracle.  The device
// arithmetic is built from single-instruction asm statements (namespace isa below).
// Semantics replaced: crates/boojum-cuda/native/goldilocks.cuh:90-183 (add/sub/mul/reduce).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gl {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u64 P = 0xFFFFFFFF00000001ull;
constexpr u64 EPS = 0xFFFFFFFFull;  // 2^64 mod p

__host__ __device__ __forceinline__ u64 canon(u64 x) { return x >= P ? x - P : x; }

// Device memory read and written through the global address space.  Through a generic pointer
// (one loaded from memory, or derived from one) the compiler emits flat_* instructions, which
// count against lgkmcnt as well as vmcnt: every wait for an LDS access then also waits for the
// flat loads and stores still in flight, serialising a tile's stores behind its LDS reads.
// global_* instructions count against vmcnt only.
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, T v) {
  *(__attribute__((address_space(1))) T *)p = v;
}

// Single gfx950 instructions, one asm statement each, with carries and lane masks in SGPR pairs:
// the compiler schedules them like any other instruction (interleaving independent work) and
// inserts the wait states a VALU-written SGPR needs before a VALU or SALU instruction reads it.
// What they cost (measured on MI355X, waves of 64 at 8 per SIMD): every instruction here issues
// at the full-width VALU rate (4.2-4.9 cycles per wave instruction per SIMD: carry-writing adds,
// v_mad_u64_u32, 64-bit compares and shifts alike), while the SALU mask logic runs beside the
// VALU; so the arithmetic below is written for the fewest VALU instructions, with the carry
// cases resolved by mask logic on the carries (plain C on the 64-bit lane masks, which the
// compiler emits as SALU instructions) and carry-in adds rather than by selects.
namespace isa {

__device__ __forceinline__ u32 lo(u64 x) { return (u32)x; }
__device__ __forceinline__ u32 hi(u64 x) { return (u32)(x >> 32); }
__device__ __forceinline__ u64 pack(u32 l, u32 h) { return ((u64)h << 32) | l; }

// a + b, carry out
__device__ __forceinline__ u32 add_co(u32 a, u32 b, u64 &co) {
  u32 r;
  asm("v_add_co_u32 %0, %1, %2, %3" : "=v"(r), "=s"(co) : "v"(a), "v"(b));
  return r;
}
// a + b + ci, carry out
__device__ __forceinline__ u32 addc(u32 a, u32 b, u64 ci, u64 &co) {
  u32 r;
  asm("v_addc_co_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(co) : "v"(a), "v"(b), "s"(ci));
  return r;
}
// a + ci, carry out
__device__ __forceinline__ u32 addc0(u32 a, u64 ci, u64 &co) {
  u32 r;
  asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(r), "=s"(co) : "v"(a), "s"(ci));
  return r;
}
__device__ __forceinline__ u32 addc0(u32 a, u64 ci) {
  u64 co;
  return addc0(a, ci, co);
}
// a - b, borrow out
__device__ __forceinline__ u32 sub_co(u32 a, u32 b, u64 &bo) {
  u32 r;
  asm("v_sub_co_u32 %0, %1, %2, %3" : "=v"(r), "=s"(bo) : "v"(a), "v"(b));
  return r;
}
// a - b - bi, borrow out
__device__ __forceinline__ u32 subb(u32 a, u32 b, u64 bi, u64 &bo) {
  u32 r;
  asm("v_subb_co_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(bo) : "v"(a), "v"(b), "s"(bi));
  return r;
}
// a - bi, borrow out
__device__ __forceinline__ u32 subb0(u32 a, u64 bi, u64 &bo) {
  u32 r;
  asm("v_subb_co_u32 %0, %1, %2, 0, %3" : "=v"(r), "=s"(bo) : "v"(a), "s"(bi));
  return r;
}
__device__ __forceinline__ u32 subb0(u32 a, u64 bi) {
  u64 bo;
  return subb0(a, bi, bo);
}
// a b + c (64-bit c), carry out
__device__ __forceinline__ u64 mad(u32 a, u32 b, u64 c, u64 &co) {
  u64 r;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(co) : "v"(a), "v"(b), "v"(c));
  return r;
}
// a b (full 64-bit product)
__device__ __forceinline__ u64 mul32(u32 a, u32 b) {
  u64 r, co;
  asm("v_mad_u64_u32 %0, %1, %2, %3, 0" : "=v"(r), "=s"(co) : "v"(a), "v"(b));
  return r;
}
// h (2^32 - 1) + l, carry out (the multiplier is the inline constant -1)
__device__ __forceinline__ u64 mad_eps(u32 h, u64 l, u64 &co) {
  u64 r;
  asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(r), "=s"(co) : "v"(h), "v"(l));
  return r;
}
// lane mask of x >= p
__device__ __forceinline__ u64 ge_p(u64 x) {
  u64 m;
  asm("v_cmp_ge_u64_e64 %0, %1, %2" : "=s"(m) : "v"(x), "s"(P));
  return m;
}
// x + EPS in the lanes of m (x + 2^32 - 1: low word - 1, high word + 1 unless the low word
// borrowed); wraps mod 2^64 where the sum overflows
__device__ __forceinline__ u64 plus_eps(u32 x0, u32 x1, u64 m) {
  u64 k;
  const u32 r0 = subb0(x0, m, k);
  return pack(r0, addc0(x1, m & ~k));
}
// x - EPS in the lanes of m (low word + 1, high word - 1 unless the low word carried)
__device__ __forceinline__ u64 minus_eps(u32 x0, u32 x1, u64 m) {
  u64 k;
  const u32 r0 = addc0(x0, m, k);
  return pack(r0, subb0(x1, m & ~k));
}

}  // namespace isa

// a, b < p: canonical a + b.  s = a + b with carry c; when c, s + EPS (= a + b - p < p); when
// not, s >= p selects s + EPS (wrapping to s - p).  5 VALU.
__device__ __forceinline__ u64 add(u64 a, u64 b) {
  using namespace isa;
  u64 c, c2;
  const u32 s0 = add_co(lo(a), lo(b), c);
  const u32 s1 = addc(hi(a), hi(b), c, c2);
  return plus_eps(s0, s1, c2 | ge_p(pack(s0, s1)));
}

// a < p, b < p: canonical a - b.  d = a - b with borrow; a borrow selects d - EPS (= a - b + p,
// in (0, p)).  With a any u64 and b < p the result is a lazy residue in [0, 2^64): d - EPS
// cannot wrap since then d >= 2^64 - b > EPS.  4 VALU.
__device__ __forceinline__ u64 sub(u64 a, u64 b) {
  using namespace isa;
  u64 c, bw;
  const u32 d0 = sub_co(lo(a), lo(b), c);
  const u32 d1 = subb(hi(a), hi(b), c, bw);
  return minus_eps(d0, d1, bw);
}

// Canonical form of any u64: x >= p selects x + EPS (wrapping to x - p).  3 VALU.
__device__ __forceinline__ u64 reduce1(u64 x) {
  using namespace isa;
  return plus_eps(lo(x), hi(x), ge_p(x));
}

__device__ __forceinline__ u64 neg(u64 a) { return a ? P - a : 0; }

// Canonical residue of lo + hl (2^32 - 1) - hh - bi (lo < 2^64; hl, hh < 2^32; bi a 0/1 lane
// mask), i.e. of lo + (hl + (hh + bi) 2^32) 2^64 by 2^64 = 2^32 - 1 and 2^96 = -1.
//   R = hl EPS + lo with carry c (one v_mad_u64_u32), D = R - hh - bi with borrow b; the value is
//   V = D + (c - b) 2^64, in (-2^33, 2^64 + 2^64), and
//     b & !c:  V < 0,           the residue is D - EPS (D >= 2^64 - 2^33, no wrap);
//     c & !b:  V > p,           the residue is D + EPS (D < hl EPS, no wrap, result < p);
//     else     V = D in [0, 2^64): D >= p selects D + EPS (wrapping to D - p).
// 8 VALU.
__device__ __forceinline__ u64 reduce3b(u64 lo_, u32 hl, u32 hh, u64 bi) {
  using namespace isa;
  u64 c, b1, b;
  const u64 R = mad_eps(hl, lo_, c);
  const u32 d0 = subb(lo(R), hh, bi, b1);
  const u32 d1 = subb0(hi(R), b1, b);
  const u64 ge = ge_p(pack(d0, d1));
  const u64 m_sub = b & ~c;
  const u64 m_add = (c & ~b) | (ge & ~(c ^ b));
  const u64 e = minus_eps(d0, d1, m_sub);
  return plus_eps(lo(e), hi(e), m_add);
}
__device__ __forceinline__ u64 reduce3(u64 lo_, u32 hl, u32 hh) {
  using namespace isa;
  u64 c, b1, b;
  const u64 R = mad_eps(hl, lo_, c);
  const u32 d0 = sub_co(lo(R), hh, b1);
  const u32 d1 = subb0(hi(R), b1, b);
  const u64 ge = ge_p(pack(d0, d1));
  const u64 m_sub = b & ~c;
  const u64 m_add = (c & ~b) | (ge & ~(c ^ b));
  const u64 e = minus_eps(d0, d1, m_sub);
  return plus_eps(lo(e), hi(e), m_add);
}

// Canonical residue of lo + hl (2^32 - 1) (lo < 2^64, hl < 2^32): R = hl EPS + lo with carry c;
// when c, R + EPS is canonical (R < hl EPS); otherwise R >= p selects R + EPS (= R - p).  4 VALU.
__device__ __forceinline__ u64 reduce2(u64 lo_, u32 hl) {
  using namespace isa;
  u64 c;
  const u64 R = mad_eps(hl, lo_, c);
  return plus_eps(lo(R), hi(R), c | ge_p(R));
}

// Reduces lo + hi * 2^64 (any 128-bit value) to a canonical residue.
__device__ __forceinline__ u64 reduce128(u64 lo_, u64 hi_) { return reduce3(lo_, (u32)hi_, (u32)(hi_ >> 32)); }

// a, b any u64; result canonical.  The 128-bit product is a0 b0 + (a0 b1 + a1 b0) 2^32 +
// a1 b1 2^64 from four v_mad_u64_u32 (the middle sum's carry cy weighs 2^96 = -1 mod p); its
// words r0..r3 reduce as [r0, r1] + r2 EPS - r3 - cy.  15 VALU, no register moves.
__device__ __forceinline__ u64 mul(u64 a, u64 b) {
  using namespace isa;
  u64 cy, k1, k2, dead;
  const u64 z = mul32(lo(a), lo(b));
  const u64 x = mul32(hi(a), lo(b));
  const u64 y = mad(lo(a), hi(b), x, cy);
  const u64 w = mul32(hi(a), hi(b));
  const u32 r1 = add_co(hi(z), lo(y), k1);
  const u32 r2 = addc(lo(w), hi(y), k1, k2);
  const u32 r3 = addc0(hi(w), k2, dead);
  return reduce3b(pack(lo(z), r1), r2, r3, cy);
}

__device__ __forceinline__ u64 sqr(u64 a) { return mul(a, a); }

// x * 2^E for 0 < E < 64, x any u64: x << E is (lo, hi) with hi < 2^E.
template <int E>
__device__ __forceinline__ u64 shl(u64 x) {
  static_assert(E > 0 && E < 64, "shift out of range");
  const u64 lo_ = x << E, hi_ = x >> (64 - E);
  if constexpr (E <= 32)
    return reduce2(lo_, (u32)hi_);
  else
    return reduce3(lo_, (u32)hi_, (u32)(hi_ >> 32));
}

// x * 2^-M for 0 < M <= 32, x any u64: with x = q 2^M + r,
// x 2^-M = q + r 2^(192-M) = q - (r 2^(32-M)) 2^64 = q - (r 2^(32-M)) (2^32 - 1),
// where q < 2^(64-M) < p and the subtrahend <= (2^32 - 1)^2 < p.
template <int M>
__device__ __forceinline__ u64 shr(u64 x) {
  static_assert(M > 0 && M <= 32, "shift out of range");
  const u64 q = x >> M;
  const u32 r = M == 32 ? (u32)x : (u32)x << (32 - M);
  u64 dead;
  return sub(q, isa::mad_eps(r, 0, dead));
}

// x * 2^E for a compile-time E in [0, 192); x any u64, result canonical.  Uses 2^96 = -1 and, for
// 64 <= E < 96, 2^E = -2^-(96-E).  `negate` reports a sign the caller must apply.
template <int E>
__device__ __forceinline__ u64 mul_pow2_signed(u64 x, bool &negate) {
  static_assert(E >= 0 && E < 192, "shift exponent out of range");
  constexpr int F = E % 96;
  constexpr bool s1 = E >= 96;
  if constexpr (F == 0) {
    negate = s1;
    return x;
  } else if constexpr (F < 64) {
    negate = s1;
    return shl<F>(x);
  } else {
    negate = !s1;
    return shr<96 - F>(x);
  }
}

template <int E>
__device__ __forceinline__ u64 mul_pow2(u64 x) {
  bool negate;
  const u64 y = mul_pow2_signed<E>(x, negate);
  return negate ? neg(y) : y;
}

// Generic runtime a^e, for tests and set-up code.
__device__ __forceinline__ u64 pow_u64(u64 a, u64 e) {
  u64 r = 1;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = sqr(a);
    e >>= 1;
  }
  return r;
}

// a^-1 = a^(p-2) (0 maps to 0)
__device__ __forceinline__ u64 inv(u64 a) { return pow_u64(a, P - 2); }

}  // namespace gl
