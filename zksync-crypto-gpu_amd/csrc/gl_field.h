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

// Reduces lo + hi * 2^64 (any 128-bit value) to a canonical residue.
__device__ __forceinline__ u64 reduce128(u64 lo, u64 hi) {
  const u32 hl = (u32)hi;
  const u32 hh = (u32)(hi >> 32);
  // lo - hh * 2^96 == lo + hh (mod p), 2^96 == -1
  u64 t0;
  if (__builtin_sub_overflow(lo, (u64)hh, &t0)) t0 -= EPS;
  // hl * 2^64 == hl * (2^32 - 1)
  const u64 t1 = ((u64)hl << 32) - hl;
  u64 r;
  if (__builtin_add_overflow(t0, t1, &r)) r += EPS;
  return canon(r);
}

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

// x * 2^E for a compile-time E in [0, 192); x canonical.
template <int E>
__device__ __forceinline__ u64 mul_pow2(u64 x) {
  static_assert(E >= 0 && E < 192, "shift exponent out of range");
  if constexpr (E == 0) {
    return x;
  } else if constexpr (E >= 96) {
    return neg(mul_pow2<E - 96>(x));  // 2^96 == -1
  } else if constexpr (E < 64) {
    return reduce128(x << E, x >> (64 - E));
  } else {
    // 64 <= E < 96, x = x1 2^32 + x0:
    //   x1 2^(E+32) = -x1 2^(E-64)      (E + 32 >= 96)
    //   x0 2^E      = z 2^64 with z = x0 << (E-64) < 2^64, z 2^64 == (z0 << 32) - z0 - z1
    constexpr int S = E - 64;
    const u64 x0 = (u32)x, x1 = x >> 32;
    const u64 z = x0 << S;
    const u64 z0 = (u32)z, z1 = z >> 32;
    const u64 y = x1 << S;  // < 2^64 since x1 < 2^32 and S < 32
    // (z0 << 32) - (z0 + z1) stays above -2^33; bring to [0, p) then subtract y
    u64 a;
    if (__builtin_sub_overflow(z0 << 32, z0 + z1, &a)) a -= EPS;
    return sub(canon(a), canon(y));
  }
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
