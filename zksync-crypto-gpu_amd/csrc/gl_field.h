// gl_field.h -- Goldilocks field arithmetic for gfx950 (p = 2^64 - 2^32 + 1).
//
// Written from the field's special form, not from the reference's carry-chain code:
//   2^64 = 2^32 - 1 (mod p),   2^96 = -1 (mod p),   2^192 = 1 (mod p).
// Every function here takes canonical operands (< p) unless it says otherwise and returns a
// canonical result, so stored results compare bit-exactly with the CPU oracle.
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

__device__ __forceinline__ u64 add(u64 a, u64 b) {
  u64 s;
  const bool carry = __builtin_add_overflow(a, b, &s);
  // s - p == s + EPS (mod 2^64); it is the answer when a + b >= p
  u64 t;
  const bool ge = __builtin_add_overflow(s, EPS, &t);
  return (carry || ge) ? t : s;
}

__device__ __forceinline__ u64 sub(u64 a, u64 b) {
  u64 d;
  const bool borrow = __builtin_sub_overflow(a, b, &d);
  return borrow ? d - EPS : d;  // d + p (mod 2^64)
}

__device__ __forceinline__ u64 neg(u64 a) { return a ? P - a : 0; }

// Canonical residue of lo + hl (2^32 - 1) - hh  (lo < 2^64, hl, hh < 2^32), which is
// lo + (hl + hh 2^32) 2^64 reduced with 2^64 = 2^32 - 1 and 2^96 = -1.
//   r = hl EPS + lo with carry c (one v_mad_u64_u32), d = r - hh with borrow b, then
//   c & !b: d + EPS (canonical, since d < hl EPS);  b & !c: d - EPS (canonical);
//   otherwise d is exact and d >= p selects d + EPS (= d - p mod 2^64).
// Written in inline asm to use the multiply-add's carry-out and to select with lane masks;
// the SALU mask logic writes SCC, hence the clobber.  Every carry a VALU instruction writes to
// an SGPR is read by a VALU instruction at least two instructions later (the VALU-writes-SGPR
// -> VALU-reads-SGPR wait states of gfx950), so the low halves of d + EPS and d - EPS are
// formed between the two halves of d.
__device__ __forceinline__ u64 reduce3(u64 lo, u32 hl, u32 hh) {
  u64 r, c;
  asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"(hl), "v"(lo));
  const u32 r0 = (u32)r, r1 = (u32)(r >> 32);
  u32 d0, d1, t0, t1, u0, u1, o0, o1;
  u64 b, ov, x, m1, m2;
  asm("v_sub_co_u32 %[d0], %[b], %[r0], %[hh]\n\t"
      "v_add_co_u32 %[t0], %[ov], %[d0], -1\n\t"
      "v_sub_co_u32 %[u0], %[x], %[d0], -1\n\t"
      "v_subb_co_u32 %[d1], %[b], %[r1], 0, %[b]\n\t"
      "v_addc_co_u32 %[t1], %[ov], %[d1], 0, %[ov]\n\t"
      "v_subb_co_u32 %[u1], %[x], %[d1], 0, %[x]\n\t"
      "s_andn2_b64 %[m1], %[c], %[b]\n\t"
      "s_xor_b64 %[m2], %[c], %[b]\n\t"
      "s_andn2_b64 %[m2], %[ov], %[m2]\n\t"
      "s_or_b64 %[m1], %[m1], %[m2]\n\t"
      "s_andn2_b64 %[m2], %[b], %[c]\n\t"
      "v_cndmask_b32 %[o0], %[d0], %[t0], %[m1]\n\t"
      "v_cndmask_b32 %[o1], %[d1], %[t1], %[m1]\n\t"
      "v_cndmask_b32 %[o0], %[o0], %[u0], %[m2]\n\t"
      "v_cndmask_b32 %[o1], %[o1], %[u1], %[m2]"
      : [d0] "=&v"(d0), [d1] "=&v"(d1), [t0] "=&v"(t0), [t1] "=&v"(t1), [u0] "=&v"(u0),
        [u1] "=&v"(u1), [o0] "=&v"(o0), [o1] "=&v"(o1), [b] "=&s"(b), [ov] "=&s"(ov),
        [x] "=&s"(x), [m1] "=&s"(m1), [m2] "=&s"(m2)
      : [r0] "v"(r0), [r1] "v"(r1), [hh] "v"(hh), [c] "s"(c)
      : "scc");
  return ((u64)o1 << 32) | o0;
}

// Canonical residue of lo + hl (2^32 - 1): r = hl EPS + lo with carry c; when c, r + EPS is
// canonical (r < hl EPS); otherwise r >= p selects r + EPS (= r - p mod 2^64).  r >= p is the
// overflow of the 64-bit sum t = r + EPS, found by comparing t < r (no carry chain).
__device__ __forceinline__ u64 reduce2(u64 lo, u32 hl) {
  u64 r, c, m;
  asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=&v"(r), "=&s"(c) : "v"(hl), "v"(lo));
  const u64 t = r + EPS;
  asm("v_cmp_lt_u64_e64 %[m], %[t], %[r]\n\t"
      "s_or_b64 %[m], %[m], %[c]"
      : [m] "=&s"(m) : [t] "v"(t), [r] "v"(r), [c] "s"(c) : "scc");
  u32 o0, o1;
  asm("v_cndmask_b32 %0, %2, %3, %6\n\t"
      "v_cndmask_b32 %1, %4, %5, %6"
      : "=&v"(o0), "=&v"(o1)
      : "v"((u32)r), "v"((u32)t), "v"((u32)(r >> 32)), "v"((u32)(t >> 32)), "s"(m));
  return ((u64)o1 << 32) | o0;
}

// Reduces lo + hi * 2^64 (any 128-bit value) to a canonical residue.
__device__ __forceinline__ u64 reduce128(u64 lo, u64 hi) { return reduce3(lo, (u32)hi, (u32)(hi >> 32)); }

// 64x64 -> 128 product from four 32x32 -> 64 multiply-adds (v_mad_u64_u32).
__device__ __forceinline__ void mul_wide(u64 a, u64 b, u64 &lo, u64 &hi) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32);
  const u32 b0 = (u32)b, b1 = (u32)(b >> 32);
  const u64 p00 = (u64)a0 * b0;
  const u64 m0 = (u64)a0 * b1 + (p00 >> 32);  // < 2^64
  const u64 m1 = (u64)a1 * b0 + (u32)m0;      // < 2^64
  hi = (u64)a1 * b1 + (m0 >> 32) + (m1 >> 32);
  lo = (m1 << 32) | (u32)p00;
}

// a, b any u64 (not necessarily canonical); result canonical.
__device__ __forceinline__ u64 mul(u64 a, u64 b) {
  u64 lo, hi;
  mul_wide(a, b, lo, hi);
  return reduce128(lo, hi);
}

__device__ __forceinline__ u64 sqr(u64 a) { return mul(a, a); }

// x * 2^E for 0 < E < 64, x canonical: x << E is (lo, hi) with hi < 2^E.
template <int E>
__device__ __forceinline__ u64 shl(u64 x) {
  static_assert(E > 0 && E < 64, "shift out of range");
  const u64 lo = x << E, hi = x >> (64 - E);
  if constexpr (E <= 32)
    return reduce2(lo, (u32)hi);
  else
    return reduce3(lo, (u32)hi, (u32)(hi >> 32));
}

// x * 2^-M for 0 < M <= 32, x canonical: with x = q 2^M + r,
// x 2^-M = q + r 2^(192-M) = q - (r 2^(32-M)) 2^64 = q - (r 2^(32-M)) (2^32 - 1),
// where q < p and the subtrahend < 2^32 (2^32 - 1) < p.
template <int M>
__device__ __forceinline__ u64 shr(u64 x) {
  static_assert(M > 0 && M <= 32, "shift out of range");
  const u64 q = x >> M;
  const u32 r = (u32)(M == 32 ? x : ((x & ((1ull << M) - 1)) << (32 - M)));
  return sub(q, (u64)r * EPS);
}

// x * 2^E for a compile-time E in [0, 192); x canonical.  Uses 2^96 = -1 and, for
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
