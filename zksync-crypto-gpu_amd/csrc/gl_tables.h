// gl_tables.h -- resident power tables and the device-side lookups that read them.
//
// A context owns, in HBM, fine/coarse power tables of omega = omega_{2^30}, g = 7 and g^-1
// (crates/boojum-cuda/src/context.rs:171-252 builds the same quantities), and 2^-k for
// k <= 30.  Unlike the reference, which publishes them through process-global __constant__
// descriptors (crates/boojum-cuda/native/context.cu:3-8), the descriptor is a plain struct
// passed by value to every kernel, so several contexts can be live at once.
#pragma once

#include "gl_field.h"

namespace gl {

constexpr u32 OMEGA_LOG_ORDER = 30;

struct PowTable {
  const u64 *fine;    // base^i for i < 2^fine_log
  const u64 *coarse;  // base^(i * 2^fine_log) for i < 2^(30 - fine_log)
  u32 fine_log;
  u32 fine_mask;
};

struct Tables {
  PowTable w;         // powers of omega_{2^30}
  PowTable g;         // powers of 7
  PowTable g_inv;     // powers of 7^-1
  const u64 *inv_sizes;  // 2^-k, k <= 30
};

// base^e for e < 2^30 (the exponent is reduced mod 2^30 first)
__device__ __forceinline__ u64 table_pow(const PowTable &t, u32 e) {
  e &= (1u << OMEGA_LOG_ORDER) - 1;
  const u64 c = gload(t.coarse + (e >> t.fine_log));
  const u32 f = e & t.fine_mask;
  return f ? mul(gload(t.fine + f), c) : c;
}

// omega_{2^30}^(+-e)
__device__ __forceinline__ u64 pow_w(const Tables &t, u32 e, bool inverse) {
  return table_pow(t.w, inverse ? (0u - e) : e);
}

}  // namespace gl
