// gl_host.h -- host-side Goldilocks arithmetic for table and launch-constant set-up.
#pragma once

#include <stdint.h>

namespace gl_host {

typedef unsigned __int128 u128;
constexpr uint64_t P = 0xFFFFFFFF00000001ull;

constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

// Canonical residue of a 128-bit value x = lo + hl 2^64 + hh 2^96 (hl, hh the 32-bit halves of
// the high word): 2^64 = EPS and 2^96 = -1 mod p, so x = lo - hh + hl EPS.  No division: called
// per launch (scale tables), where a 128-bit % costs tens of nanoseconds.
inline uint64_t reduce128(u128 x) {
  const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  const uint64_t hl = hi & EPS, hh = hi >> 32;
  uint64_t t = lo - hh;
  if (lo < hh) t -= EPS;  // borrowed 2^64 = EPS: t = lo - hh + 2^64 >= 2^64 - 2^32 > EPS
  const uint64_t u = hl * EPS;
  uint64_t r = t + u;
  if (r < u) r += EPS;  // carried 2^64: r + EPS < 2^64 since t + u < 2^64 + 2^64 - 2^33
  return r >= P ? r - P : r;
}

inline uint64_t mul(uint64_t a, uint64_t b) { return reduce128((u128)a * b); }

inline uint64_t pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  a = a >= P ? a - P : a;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return r;
}

inline uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// omega_{2^k}: the 2^32-th root 7^((p-1)/2^32) squared (32 - k) times
inline uint64_t omega(unsigned log_size) {
  uint64_t w = pow(7, (P - 1) >> 32);
  for (unsigned i = log_size; i < 32; i++) w = mul(w, w);
  return w;
}

// coset shift s = 7 omega_{2^(log_n + log_ext)}^coset_idx, or 1 without extension
inline uint64_t coset_shift(unsigned log_n, unsigned log_ext, unsigned coset_idx) {
  if (log_ext == 0) return 1;
  return mul(7, pow(omega(log_n + log_ext), coset_idx));
}

}  // namespace gl_host
